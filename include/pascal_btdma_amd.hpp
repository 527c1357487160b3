// Header-only C++ mirror of the reference's C++ interface
// (CUDA_CXX_Port/include/pascal_btdma_cuda.hpp:11-168) on top of the C ABI in
// pascal_btdma_amd.h: same namespace, class and function names, argument meaning and
// exception types, so a caller of the reference port changes its include and replaces
// the MPI communicator by the number of row partitions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>

#include "pascal_btdma_amd.h"

namespace pascal_btdma {

constexpr int kMaxBlockSize = BTD_MAX_M;

// The reference's CudaError (pascal_btdma_cuda.hpp:20-22).
struct HipError : public std::runtime_error {
    explicit HipError(const std::string& message) : std::runtime_error(message) {}
};

// A zero pivot met in some block LU (no counterpart: the reference ignores it).
struct SingularError : public std::runtime_error {
    explicit SingularError(const std::string& message) : std::runtime_error(message) {}
};

// pascal_btdma_cuda.hpp:24-39.
struct BtdmaSolveTimings {
    double total = 0.0;
    double local_compute = 0.0;
    double forward_exchange = 0.0;
    double reduced_compute = 0.0;
    double backward_exchange = 0.0;
    double update_compute = 0.0;

    double computation() const { return local_compute + reduced_compute + update_compute; }
    double communication() const { return forward_exchange + backward_exchange; }
};

inline void check_status(btd_status s) {
    if (s == BTD_OK) return;
    const std::string msg = btd_last_error();
    switch (s) {
        case BTD_ERR_INVALID_ARG:
        case BTD_ERR_DIM_MISMATCH: throw std::invalid_argument(msg);
        case BTD_ERR_NOT_CREATED: throw std::runtime_error(msg);
        case BTD_ERR_HIP: throw HipError(msg);
        case BTD_ERR_SINGULAR: throw SingularError(msg);
        case BTD_ERR_NOT_IMPLEMENTED: throw std::logic_error(msg);
        default: throw std::runtime_error("pascal_btdma_amd: status " + std::to_string((int)s));
    }
}

inline std::size_t matrix_index(int sys, int row, int p, int q, int nsys, int nrow, int m) {
    return btd_matrix_index(sys, row, p, q, nsys, nrow, m);
}

inline std::size_t vector_index(int sys, int row, int p, int nsys, int nrow) {
    return btd_vector_index(sys, row, p, nsys, nrow);
}

inline void partition_1d(int start, int end, int nprocs, int rank, int& first, int& last) {
    btd_partition_1d(start, end, nprocs, rank, &first, &last);
}

// pascal_btdma_cuda.hpp:88-135; npart replaces the MPI communicator.
struct BtdmaGpuPlan {
    btd_plan_t handle = nullptr;
    int m = 0, nsys = 0, nrow = 0, npart = 1;
    btd_layout layout = BTD_LAYOUT_REF_GPU;
    bool created = false;

    BtdmaGpuPlan() = default;
    BtdmaGpuPlan(const BtdmaGpuPlan&) = delete;
    BtdmaGpuPlan& operator=(const BtdmaGpuPlan&) = delete;
    BtdmaGpuPlan(BtdmaGpuPlan&& o) noexcept { swap(o); }
    BtdmaGpuPlan& operator=(BtdmaGpuPlan&& o) noexcept {
        if (this != &o) {
            destroy();
            swap(o);
        }
        return *this;
    }
    ~BtdmaGpuPlan() { destroy(); }

    void create(int m_in, int nsys_in, int nrow_in, int npart_in = 1,
                btd_layout layout_in = BTD_LAYOUT_REF_GPU) {
        destroy();
        check_status(btd_plan_create(&handle, m_in, nsys_in, nrow_in, npart_in, layout_in));
        m = m_in;
        nsys = nsys_in;
        nrow = nrow_in;
        npart = npart_in;
        if (npart == 0) check_status(btd_plan_npart(handle, &npart));  // the library's choice
        layout = layout_in;
        created = true;
    }

    void destroy() noexcept {
        if (handle) btd_plan_destroy(handle);
        handle = nullptr;
        created = false;
    }

    // Synchronises; true if a zero pivot was met since the last check.
    bool singular() {
        int s = 0;
        check_status(btd_plan_check(handle, &s));
        return s != 0;
    }

    // Synchronises; info[s] = 1 for every system that met a zero pivot since the last check
    // (info must hold nsys ints).
    void singular_systems(int* info) { check_status(btd_plan_info(handle, info)); }

    // How an npart = 1 plan's non-cyclic solve runs (btd_plan_set_form): BTD_FORM_PLAIN, the
    // library's choice (the default), BTD_FORM_TWISTED, always from both ends,
    // BTD_FORM_LAPACK_ORDER, one lane per system in the CPU path's BLAS/LAPACK operation order,
    // or BTD_FORM_WORKSPACE, that solve with C' in a plan-owned workspace (C only read).
    void set_form(btd_form f) { check_status(btd_plan_set_form(handle, (int)f)); }

    // BTD_FORM_WORKSPACE only: keep the fastest of up to `candidates` workspace allocations for
    // these arrays (btd_plan_place_workspace); ms: NULL or `candidates` entries.
    void place_workspace(const double* a, const double* b, const double* c, double* d, int candidates,
                         double* ms = nullptr, hipStream_t stream = nullptr) {
        check_status(btd_plan_place_workspace(handle, a, b, c, d, m, nsys, nrow, candidates, ms, stream));
    }
    // The same for the FP32-storage solve (solve_noncyclic_f32s with work_dev == nullptr).
    void place_workspace_f32s(const float* a, const float* b, const float* c, double* d, int candidates,
                              double* ms = nullptr, hipStream_t stream = nullptr) {
        check_status(btd_plan_place_workspace_f32s(handle, a, b, c, d, candidates, ms, stream));
    }
    btd_form form() {
        int f = 0;
        check_status(btd_plan_form(handle, &f));
        return (btd_form)f;
    }

   private:
    void swap(BtdmaGpuPlan& o) noexcept {
        std::swap(handle, o.handle);
        std::swap(m, o.m);
        std::swap(nsys, o.nsys);
        std::swap(nrow, o.nrow);
        std::swap(npart, o.npart);
        std::swap(layout, o.layout);
        std::swap(created, o.created);
    }
};

inline void solve_noncyclic(BtdmaGpuPlan& plan, double* a_dev, double* b_dev, double* c_dev,
                            double* d_dev, int m, int nsys, int nrow, hipStream_t stream = nullptr) {
    check_status(btd_solve_noncyclic(plan.handle, a_dev, b_dev, c_dev, d_dev, m, nsys, nrow,
                                     (void*)stream));
}

inline void solve_noncyclic_profiled(BtdmaGpuPlan& plan, double* a_dev, double* b_dev,
                                     double* c_dev, double* d_dev, int m, int nsys, int nrow,
                                     BtdmaSolveTimings* timings, hipStream_t stream = nullptr) {
    btd_timings t{};
    check_status(btd_solve_noncyclic_profiled(plan.handle, a_dev, b_dev, c_dev, d_dev, m, nsys,
                                              nrow, &t, (void*)stream));
    if (timings) {
        timings->total = t.total;
        timings->local_compute = t.local_compute;
        timings->forward_exchange = t.forward_exchange;
        timings->reduced_compute = t.reduced_compute;
        timings->backward_exchange = t.backward_exchange;
        timings->update_compute = t.update_compute;
    }
}

inline void solve_cyclic(BtdmaGpuPlan& plan, double* a_dev, double* b_dev, double* c_dev,
                         double* d_dev, int m, int nsys, int nrow, hipStream_t stream = nullptr) {
    check_status(btd_solve_cyclic(plan.handle, a_dev, b_dev, c_dev, d_dev, m, nsys, nrow,
                                  (void*)stream));
}

// Entry points with no counterpart in the reference (see the C header for their contracts).

// FP32-stored blocks, FP64 arithmetic; work: btd_layout_elems(layout, nsys, nrow, m*m) doubles, or
// nullptr for the workspace of a plan in BTD_FORM_WORKSPACE.
inline void solve_noncyclic_f32s(BtdmaGpuPlan& plan, const float* a_dev, const float* b_dev,
                                 const float* c_dev, double* d_dev, double* work_dev,
                                 hipStream_t stream = nullptr) {
    check_status(btd_solve_noncyclic_f32s(plan.handle, a_dev, b_dev, c_dev, d_dev, work_dev,
                                          (void*)stream));
}

// euler3d sweeps: axis 0/1/2 = x/y/z; U(0:n1+1,0:n2+1,0:n3+1,5), fields (n1,n2,n3,5).
inline void euler_assemble(BtdmaGpuPlan& plan, int axis, const double* U_dev, const double* rhs_dev,
                           int n1, int n2, int n3, double h, double dt, double eps, double* a_dev,
                           double* b_dev, double* c_dev, double* d_dev, hipStream_t stream = nullptr) {
    check_status(btd_euler_assemble(plan.handle, axis, U_dev, rhs_dev, n1, n2, n3, h, dt, eps, a_dev,
                                    b_dev, c_dev, d_dev, (void*)stream));
}

inline void sweep_scatter(BtdmaGpuPlan& plan, int axis, const double* d_dev, int n1, int n2, int n3,
                          double* field_dev, hipStream_t stream = nullptr) {
    check_status(btd_sweep_scatter(plan.handle, axis, d_dev, n1, n2, n3, field_dev, (void*)stream));
}

inline void euler_sweep_cyclic(BtdmaGpuPlan& plan, int axis, const double* U_dev,
                               const double* rhs_dev, int n1, int n2, int n3, double h, double dt,
                               double eps, double* work_dev, double* out_dev,
                               hipStream_t stream = nullptr) {
    check_status(btd_euler_sweep_cyclic(plan.handle, axis, U_dev, rhs_dev, n1, n2, n3, h, dt, eps,
                                        work_dev, out_dev, (void*)stream));
}

}  // namespace pascal_btdma
