// The reference's C++ example (CUDA_CXX_Port/examples/ex_btdma_noncyclic.cu) ported to the
// C++ mirror header: the Study's diagonal workload (A = 0.25 I, B = 2 I, C = 0.25 I, d = 1),
// 32 x 32 systems of 128 block rows, solved with npart = 1, npart = 4, the partition count
// the library chooses (npart = 0) and npart = 1 in the twisted form (plan.set_form), checked against
// the reference's published signature (sum, L2, end and middle samples) and against the
// error behaviour of the reference (invalid sizes, dimension mismatch, uncreated plan).
// Usage: ex_btdma_noncyclic <m> <expected_sum> <expected_l2>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "../../include/pascal_btdma_amd.hpp"

#define HIPCK(x)                                                               \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) {                                                \
            std::fprintf(stderr, "HIP error %s: %s\n", #x, hipGetErrorString(e_)); \
            return 2;                                                          \
        }                                                                      \
    } while (0)

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s m expected_sum expected_l2\n", argv[0]);
        return 2;
    }
    const int m = std::atoi(argv[1]);
    const double want_sum = std::atof(argv[2]), want_l2 = std::atof(argv[3]);
    const int n1 = 32, n2 = 32, n3 = 128, nsys = n1 * n2, nrow = n3;
    const size_t ms = (size_t)nsys * nrow * m * m, vs = (size_t)nsys * nrow * m;
    std::vector<double> ha(ms, 0.0), hb(ms, 0.0), hc(ms, 0.0), hd(vs, 1.0);
    for (int row = 0; row < nrow; ++row)
        for (int sys = 0; sys < nsys; ++sys)
            for (int p = 0; p < m; ++p) {
                ha[pascal_btdma::matrix_index(sys, row, p, p, nsys, nrow, m)] = 0.25;
                hb[pascal_btdma::matrix_index(sys, row, p, p, nsys, nrow, m)] = 2.00;
                hc[pascal_btdma::matrix_index(sys, row, p, p, nsys, nrow, m)] = 0.25;
            }
    double *a, *b, *c, *d;
    HIPCK(hipMalloc(&a, ms * 8));
    HIPCK(hipMalloc(&b, ms * 8));
    HIPCK(hipMalloc(&c, ms * 8));
    HIPCK(hipMalloc(&d, vs * 8));
    int failures = 0;
    for (int npart : {1, 4, 0, -1}) {  // -1: npart = 1, always twisted (plan.set_form)
        HIPCK(hipMemcpy(a, ha.data(), ms * 8, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(b, hb.data(), ms * 8, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(c, hc.data(), ms * 8, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(d, hd.data(), vs * 8, hipMemcpyHostToDevice));
        pascal_btdma::BtdmaGpuPlan plan;
        plan.create(m, nsys, nrow, npart < 0 ? 1 : npart);
        if (npart < 0) {
            plan.set_form(BTD_FORM_TWISTED);
            if (plan.form() != BTD_FORM_TWISTED) {
                std::printf("set_form: the plan's form did not change\n");
                ++failures;
            }
        }
        if (npart == 0 && plan.npart != btd_auto_npart(m, nsys, nrow)) {
            std::printf("npart=0: plan.npart %d, btd_auto_npart %d\n", plan.npart, btd_auto_npart(m, nsys, nrow));
            ++failures;
        }
        pascal_btdma::BtdmaSolveTimings t;
        pascal_btdma::solve_noncyclic_profiled(plan, a, b, c, d, m, nsys, nrow, &t);
        std::vector<double> x(vs);
        HIPCK(hipMemcpy(x.data(), d, vs * 8, hipMemcpyDeviceToHost));
        double s = 0, s2 = 0;
        for (double v : x) {
            s += v;
            s2 += v * v;
        }
        const double x0 = x[pascal_btdma::vector_index(0, 0, 0, nsys, nrow)];
        const double xm = x[pascal_btdma::vector_index(0, n3 / 2, 0, nsys, nrow)];
        const bool ok = std::fabs(s / want_sum - 1) < 1e-9 && std::fabs(std::sqrt(s2) / want_l2 - 1) < 1e-9 &&
                        std::fabs(x0 - 0.4508066615170332) < 1e-14 && std::fabs(xm - 0.4) < 1e-14 &&
                        t.communication() == 0.0 && t.computation() > 0.0;
        std::printf("m=%d npart=%d%s sum=%.16g l2=%.16g x0=%.16g xmid=%.16g total=%.3es %s\n", m, plan.npart,
                    npart < 0 ? " (twisted)" : "", s, std::sqrt(s2), x0, xm, t.total, ok ? "OK" : "MISMATCH");
        failures += ok ? 0 : 1;
    }
    // error behaviour of the reference (pascal_btdma_cuda.cu:865-870,962-967,1063-1075)
    auto expect = [&](const char* what, auto fn, auto tag) {
        try {
            fn();
            std::printf("%s: no exception\n", what);
            ++failures;
        } catch (const decltype(tag)&) {
            std::printf("%s: raised as expected\n", what);
        } catch (...) {
            std::printf("%s: wrong exception type\n", what);
            ++failures;
        }
    };
    expect("m = 11", [&] { pascal_btdma::BtdmaGpuPlan p; p.create(11, nsys, nrow); }, std::invalid_argument(""));
    expect("nrow = 1", [&] { pascal_btdma::BtdmaGpuPlan p; p.create(m, nsys, 1); }, std::invalid_argument(""));
    expect("uncreated plan", [&] { pascal_btdma::BtdmaGpuPlan p; pascal_btdma::solve_noncyclic(p, a, b, c, d, m, nsys, nrow); },
           std::runtime_error(""));
    expect("dimension mismatch", [&] { pascal_btdma::BtdmaGpuPlan p; p.create(m, nsys, nrow);
                                       pascal_btdma::solve_noncyclic(p, a, b, c, d, m, nsys, nrow - 1); },
           std::invalid_argument(""));
    expect("cyclic with nrow = 2", [&] { pascal_btdma::BtdmaGpuPlan p; p.create(m, nsys, 2);
                                         pascal_btdma::solve_cyclic(p, a, b, c, d, m, nsys, 2); },
           std::invalid_argument(""));
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(c);
    (void)hipFree(d);
    std::printf("%s\n", failures ? "FAILED" : "PASSED");
    return failures ? 1 : 0;
}
