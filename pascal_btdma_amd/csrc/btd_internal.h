// Internal declarations shared by the kernel and host translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "btd_common.h"
#include "btd_dense.h"

namespace btd {

hipError_t launch_plain(int m, int layout, const double* A, const double* B, double* C, double* D,
                        int nsys, int nrow, int* flag, hipStream_t s);

// The twisted cooperative single-partition solve (m >= 2), either layout.
hipError_t launch_plain_twisted(int m, int layout, const double* A, const double* B, double* C, double* D,
                                int nsys, int nrow, int* flag, hipStream_t s);

hipError_t launch_plain_save(int m, int layout, const double* A, const double* B, const double* C,
                             double* Cw, const double* Din, double* D, double* F, uint64_t* PV, int nsys,
                             int nrow, int* flag, hipStream_t s);

hipError_t launch_adjoint(int m, int layout, const double* A, const double* Cp, const double* F,
                          const uint64_t* PV, const double* X, const double* W, double* Lam,
                          double* dA, double* dB, double* dC, int nsys, int nrow, hipStream_t s);

hipError_t launch_f32r(int m, int layout, const float* A, const float* B, const float* C,
                       const double* D, double* X, float* Wc, float* Wv, int nsys, int nrow,
                       int sweeps, int* flag, hipStream_t s, bool transposed);

hipError_t launch_euler_assemble(int layout, int axis, const double* U, const double* F, int n1,
                                 int n2, int n3, double cc, double cd, double* A, double* B,
                                 double* C, double* D, int nsys, int nrow, hipStream_t s);

// weno2d sweep blocks (m = 3) from the haloed scalar field and each line's speed (btd_weno.hip).
hipError_t launch_weno_assemble(int layout, int axis, const double* U, const double* vel, int n1, int n2,
                                double h, double gdt, double* A, double* B, double* C, int nsys, int nrow,
                                hipStream_t s);
// Matrix-free periodic weno2d sweep: the two-stage periodic solve with stage 1 computing the
// blocks from U (btd_weno.hip); Aw, Cw hold stage 1's output over the plan's matrix size.
hipError_t launch_weno_sweep_cyclic(int layout, int axis, const double* U, const double* vel, int n1, int n2,
                                    double h, double gdt, double* Aw, double* Cw, double* D, int nsys, int nrow,
                                    const int* pf, const int* pl, int npart, double* redA, double* redC,
                                    double* redD, int* flag, hipStream_t s);

// weno2d scalar field F(n1, n2) <-> sweep vector (m = 3) in the plan's layout (btd_weno.hip).
hipError_t launch_weno_move(int layout, int axis, bool to_d, double* F, double* D, int n1, int nsys, int nrow,
                            hipStream_t s);

hipError_t launch_euler_sweep_cyclic(int layout, int axis, const double* U, const double* rhs,
                                     int n1, int n2, int n3, double cc, double cd, double* work,
                                     double* F, int nsys, int nrow, int* flag, hipStream_t s);

hipError_t launch_sweep_move(int layout, int axis, double* F, double* D, int n1, int n2, int n3,
                             int m, int nsys, int nrow, bool scatter, hipStream_t s);

hipError_t launch_plain_f32s(int m, int layout, const float* A, const float* B, const float* C,
                             double* Cw, double* D, int nsys, int nrow, int* flag, hipStream_t s);

// The single-partition solve one lane per system (the oracle's operation order) at any size,
// C' written to Cw (Cw = C: in place).
hipError_t launch_plain_ref(int m, int layout, const double* A, const double* B, const double* C, double* Cw,
                            double* D, int nsys, int nrow, int* flag, hipStream_t s);

hipError_t launch_outer_f32(int m, int layout, const double* Lam, const double* X, float* dA,
                            float* dB, float* dC, int nsys, int nrow, hipStream_t s);

hipError_t launch_two_stage(int m, int layout, double* A, const double* B, double* C, double* D,
                            int nsys, int nrow, const int* part_first, const int* part_len,
                            int npart, double* redA, double* redC, double* redD, int* flag,
                            hipStream_t s, hipEvent_t* ev, bool cyclic);

hipError_t launch_cyclic(int m, int layout, double* A, const double* B, double* C, double* D,
                         int nsys, int nrow, int* flag, hipStream_t s);

hipError_t launch_generate(int m, int kind, uint64_t seed, int layout, int nsys, int nrow, double* A,
                           double* B, double* C, double* D, float* Af, float* Bf, float* Cf,
                           hipStream_t s);

// Device tables of a packed mixed-size batch (see btd_batch.h).
struct BatchDev {
    const int* item_pack;      // work items in work order: first pack,
    const int* item_npack;     // number of packs (consecutive, same m),
    const int* item_mode;      // 0 lane per system (VGPR), 1 lane per system (LDS),
                               // 2 m lanes per system (one block row each), 3 two groups
                               // of m lanes per system (twisted)
    const int* pack_m;         // block size of each pack
    const int* pack_rows;      // rows of the longest system of each pack (its slot 0)
    const int* pack_count;     // systems in the pack (<= batch_width(m))
    const int* pack_slot0;     // first entry of the pack in slot_rows
    const int* slot_rows;      // rows of each system slot of each pack
    const int64_t* mat_off;    // first matrix element of each pack
    const int64_t* vec_off;    // first vector element of each pack
    const int* nitems;         // number of work items (device, written with the item table)
    int npacks;
};

hipError_t launch_batch(const BatchDev& b, int kind, int grid, const double* A, const double* B,
                        double* C, double* D, int* counter, int* flag, hipStream_t s);

hipError_t launch_batch_generate(const BatchDev& b, const int* slot_sys, int max_rows, int kind,
                                 uint64_t seed, double* A, double* B, double* C, double* D,
                                 hipStream_t s);

// Sets the thread-local message returned by btd_last_error().
void set_last_error(const char* msg);

// Compute units of the current device (queried once per device).
int device_cus();

hipError_t launch_convert(int src_layout, int dst_layout, int E, int nsys, int nrow,
                          const double* src, double* dst, hipStream_t s);

}  // namespace btd
