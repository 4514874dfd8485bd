// C++ test of the linalg_backend::ROCM_NATIVE specialisations (include/linalg_rocm_native.hpp),
// in the style of the reference's backend tests (TestGpuMatrixOperations.cxx: device buffers,
// workspace from workspace_size(), throws on invalid arguments, info on singular input).
//
//   test_rocm_native_backend --no-gpu   argument checks and workspace sizes only (no HIP call)
//   test_rocm_native_backend            also LU + solve on the GPU, checked against the CPU
//                                        oracle (identical pivots, LU and X within 1e-12
//                                        relative), the zero-pivot KAT and a mixed solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "linalg_rocm_native.hpp"

extern "C" {
int ora_gesv_d(double* m, long long lda, long long* pivot, double* b, long long nrhs, long long ldb);
}

using linalg::linalg_backend;
using linalg::linear_solver_type;
using LU = linalg::lu_decomposition<double, linalg_backend::ROCM_NATIVE>;
using SV = linalg::linear_solver<double, linalg_backend::ROCM_NATIVE>;

static int g_fail = 0;
#define EXPECT(cond)                                                         \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++g_fail;                                                        \
        }                                                                    \
    } while (0)

template <typename E, typename F> static bool throws(F f)
{
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

// SplitMix64 -> uniform (0, 1] -> Box-Muller: a portable N(0, 1) stream
struct Normal {
    uint64_t s;
    bool have = false;
    double spare = 0.0;
    explicit Normal(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (static_cast<double>(next() >> 11) + 1.0) * 0x1.0p-53; }
    double operator()()
    {
        if (have) {
            have = false;
            return spare;
        }
        const double r = std::sqrt(-2.0 * std::log(uniform())), t = 2.0 * M_PI * uniform();
        spare = r * std::sin(t);
        have = true;
        return r * std::cos(t);
    }
};

static void host_checks()
{
    EXPECT(LU::workspace_size(1024) > 0);
    EXPECT(SV::workspace_size(1024, linear_solver_type::LU_LINEAR_SOLVER) >=
           SV::workspace_size(1024, linear_solver_type::LU_UPFRONT_LINEAR_SOLVER));
    long long piv[4];
    double m[16];
    double ws[8];
    EXPECT(throws<std::invalid_argument>([&] { LU::apply(nullptr, 4, piv, ws, sizeof(ws)); }));
    EXPECT(throws<std::invalid_argument>([&] { LU::apply(m, 4, nullptr, ws, sizeof(ws)); }));
    EXPECT(throws<std::invalid_argument>([&] { LU::apply(m, 0, piv, ws, sizeof(ws)); }));
    EXPECT(throws<std::invalid_argument>([&] { LU::apply(m, 4, piv, ws, 1); }));  // workspace too small
    EXPECT(throws<std::invalid_argument>(
        [&] { SV::apply(nullptr, piv, 4, m, linear_solver_type::LU_LINEAR_SOLVER, ws, sizeof(ws)); }));
    EXPECT(throws<std::invalid_argument>(
        [&] { SV::apply(m, piv, 4, m, linear_solver_type::CHOLESKY_LINEAR_SOLVER, ws, sizeof(ws)); }));
    EXPECT(SV::workspace_size(4, linear_solver_type::CHOLESKY_LINEAR_SOLVER) > 0);
    EXPECT(throws<std::invalid_argument>([&] { SV::workspace_size(4, static_cast<linear_solver_type>(5)); }));
    // the triangular solver's argument checks (dimensions, aliasing, strides) run on the host
    using TS = linalg::triangular_solver<double, linalg::linalg_backend::ROCM_NATIVE>;
    double tb[16], tx[16];
    EXPECT(TS::workspace_size(70, 3) > 0);
    EXPECT(throws<std::invalid_argument>([&] { TS::workspace_size(0, 2); }));
    EXPECT(throws<std::invalid_argument>([&] { TS::workspace_size(2, -1); }));
    EXPECT(throws<std::invalid_argument>([&] { TS::apply(true, false, false, 4, 2, m, 4, tb, 2, tb, 2, ws, sizeof(ws)); }));
    EXPECT(throws<std::invalid_argument>([&] { TS::apply(true, false, false, 4, 2, m, 3, tb, 2, tx, 2, ws, sizeof(ws)); }));
    // determinant / inversion: argument checks on the host, Cholesky modes not provided
    using DT = linalg::determinant<double, linalg::linalg_backend::ROCM_NATIVE>;
    using IV = linalg::inversion<double, linalg::linalg_backend::ROCM_NATIVE>;
    EXPECT(DT::workspace_size(100, linear_solver_type::LU_LINEAR_SOLVER) >
           DT::workspace_size(100, linear_solver_type::LU_UPFRONT_LINEAR_SOLVER));
    EXPECT(IV::workspace_size(100, linear_solver_type::LU_LINEAR_SOLVER) >= 100 * 100 * sizeof(double));
    EXPECT(throws<std::invalid_argument>([&] { DT::apply(nullptr, nullptr, 4, linear_solver_type::LU_LINEAR_SOLVER, ws, sizeof(ws)); }));
    EXPECT(throws<std::invalid_argument>([&] { DT::apply(m, nullptr, 0, linear_solver_type::LU_LINEAR_SOLVER, ws, sizeof(ws)); }));
    EXPECT(DT::workspace_size(4, linear_solver_type::CHOLESKY_LINEAR_SOLVER) > 0);
    EXPECT(throws<std::invalid_argument>([&] { DT::workspace_size(4, static_cast<linear_solver_type>(5)); }));
    EXPECT(throws<std::invalid_argument>([&] { IV::apply(nullptr, nullptr, 4, linear_solver_type::LU_LINEAR_SOLVER, ws, sizeof(ws)); }));
    EXPECT(throws<std::invalid_argument>([&] { IV::apply(m, nullptr, 4, static_cast<linear_solver_type>(5), ws, sizeof(ws)); }));
    // cholesky_decomposition: device info required (the reference's GPU contract), triangle checked
    using CH = linalg::cholesky_decomposition<double, linalg::linalg_backend::ROCM_NATIVE>;
    int dinfo = 0;
    EXPECT(CH::workspace_size(100, linalg::cholesky_decomposition_enum::LOWER_TRIANGULAR) > 0);
    EXPECT(CH::workspace_size(100, linalg::cholesky_decomposition_enum::UPPER_TRIANGULAR) ==
           CH::workspace_size(100, linalg::cholesky_decomposition_enum::LOWER_TRIANGULAR));
    EXPECT(throws<std::invalid_argument>(
        [&] { CH::apply(m, 4, linalg::cholesky_decomposition_enum::LOWER_TRIANGULAR, ws, sizeof(ws), nullptr); }));
    EXPECT(throws<std::invalid_argument>(
        [&] { CH::apply(nullptr, 4, linalg::cholesky_decomposition_enum::LOWER_TRIANGULAR, ws, sizeof(ws), &dinfo); }));
    EXPECT(throws<std::invalid_argument>(
        [&] { CH::apply(m, 0, linalg::cholesky_decomposition_enum::LOWER_TRIANGULAR, ws, sizeof(ws), &dinfo); }));
    EXPECT(throws<std::invalid_argument>(
        [&] { CH::apply(m, 4, static_cast<linalg::cholesky_decomposition_enum>('X'), ws, sizeof(ws), &dinfo); }));
    EXPECT(throws<std::invalid_argument>(
        [&] { CH::apply(m, 4, linalg::cholesky_decomposition_enum::LOWER_TRIANGULAR, ws, 1, &dinfo); }));
    // batched multiplication: the reference's checks (pointers, then sizes) before any launch
    EXPECT(throws<std::invalid_argument>(
        [&] { linalg::rocm_native::batched_matrix_multiplication(2, 2, static_cast<const double*>(nullptr), m, m); }));
    EXPECT(throws<std::invalid_argument>([&] { linalg::rocm_native::batched_matrix_multiplication(0, 2, m, m, m); }));
}

#define HIP_OK(x) EXPECT((x) == hipSuccess)

static double rel_diff(const std::vector<double>& a, const std::vector<double>& b)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        num = std::fmax(num, std::fabs(a[i] - b[i]));
        den = std::fmax(den, std::fabs(b[i]));
    }
    return den > 0.0 ? num / den : num;
}

static void gpu_checks()
{
    const long long n = 300;
    Normal rng(20241203);
    std::vector<double> a(n * n), x(n);
    for (auto& v : a) v = rng();
    for (auto& v : x) v = rng();

    // oracle
    std::vector<double> lu_ref = a, x_ref = x;
    std::vector<long long> piv_ref(n);
    EXPECT(ora_gesv_d(lu_ref.data(), n, piv_ref.data(), x_ref.data(), 1, 1) == 0);

    double *dm = nullptr, *dx = nullptr;
    long long* dp = nullptr;
    void* dws = nullptr;
    const size_t ws = SV::workspace_size(n, linear_solver_type::LU_LINEAR_SOLVER);
    HIP_OK(hipMalloc(&dm, n * n * sizeof(double)));
    HIP_OK(hipMalloc(&dx, n * sizeof(double)));
    HIP_OK(hipMalloc(&dp, n * sizeof(long long)));
    HIP_OK(hipMalloc(&dws, ws));
    HIP_OK(hipMemcpy(dm, a.data(), n * n * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dx, x.data(), n * sizeof(double), hipMemcpyHostToDevice));
    SV::apply(dm, dp, n, dx, linear_solver_type::LU_LINEAR_SOLVER, dws, ws);
    HIP_OK(hipDeviceSynchronize());
    std::vector<double> lu(n * n), xs(n);
    std::vector<long long> piv(n);
    HIP_OK(hipMemcpy(lu.data(), dm, n * n * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(xs.data(), dx, n * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(piv.data(), dp, n * sizeof(long long), hipMemcpyDeviceToHost));
    EXPECT(piv == piv_ref);
    const double e_lu = rel_diff(lu, lu_ref), e_x = rel_diff(xs, x_ref);
    std::printf("n=%lld LU rel diff %.2e, X rel diff %.2e\n", n, e_lu, e_x);
    EXPECT(e_lu <= 1e-12);
    EXPECT(e_x <= 1e-12);

    // LU_UPFRONT reuses the factors and pivots
    HIP_OK(hipMemcpy(dx, x.data(), n * sizeof(double), hipMemcpyHostToDevice));
    SV::apply(dm, dp, n, dx, linear_solver_type::LU_UPFRONT_LINEAR_SOLVER, dws, ws);
    HIP_OK(hipMemcpy(xs.data(), dx, n * sizeof(double), hipMemcpyDeviceToHost));
    EXPECT(rel_diff(xs, x_ref) <= 1e-12);

    // zero-pivot KAT (TestLUDecomposition.cxx:283-299 style): a zero first column -> info 1
    std::vector<double> z(4 * 4, 0.0);
    for (int i = 0; i < 4; ++i)
        for (int j = 1; j < 4; ++j) z[i * 4 + j] = 1.0 + i + 2.0 * j;
    HIP_OK(hipMemcpy(dm, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice));
    int info = -1;
    const bool ok = LU::apply(dm, 4, dp, dws, LU::workspace_size(4), &info);
    EXPECT(!ok);
    EXPECT(info == 1);

    // mixed-precision solve of a diagonally dominant system
    const long long nm = 512;
    std::vector<double> am(nm * nm), bm(nm);
    for (long long i = 0; i < nm; ++i) {
        for (long long j = 0; j < nm; ++j) am[i * nm + j] = rng();
        am[i * nm + i] += static_cast<double>(nm);
        bm[i] = rng();
    }
    double *da = nullptr, *db = nullptr, *dxm = nullptr;
    void* dwm = nullptr;
    const size_t wsm = linalg::rocm_native::gesv_mixed_workspace_size(nm);
    HIP_OK(hipMalloc(&da, nm * nm * sizeof(double)));
    HIP_OK(hipMalloc(&db, nm * sizeof(double)));
    HIP_OK(hipMalloc(&dxm, nm * sizeof(double)));
    HIP_OK(hipMalloc(&dwm, wsm));
    HIP_OK(hipMemcpy(da, am.data(), am.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(db, bm.data(), bm.size() * sizeof(double), hipMemcpyHostToDevice));
    const auto r = linalg::rocm_native::gesv_mixed(da, nm, db, dxm, dwm, wsm);
    std::vector<double> xm(nm);
    HIP_OK(hipMemcpy(xm.data(), dxm, nm * sizeof(double), hipMemcpyDeviceToHost));
    double rn = 0.0, an = 0.0, xn = 0.0, bn = 0.0;
    for (long long i = 0; i < nm; ++i) {
        double s = bm[i], rs = 0.0;
        for (long long j = 0; j < nm; ++j) {
            s -= am[i * nm + j] * xm[j];
            rs += std::fabs(am[i * nm + j]);
        }
        rn = std::fmax(rn, std::fabs(s));
        an = std::fmax(an, rs);
        xn = std::fmax(xn, std::fabs(xm[i]));
        bn = std::fmax(bn, std::fabs(bm[i]));
    }
    const double berr = rn / (an * xn + bn);
    std::printf("mixed n=%lld path=%d ir=%d backward error %.2e\n", nm, r.path, r.ir_iterations, berr);
    EXPECT(berr <= 1e-14);
    EXPECT(r.path == 0);
    (void)hipFree(dm);
    (void)hipFree(dx);
    (void)hipFree(dp);
    (void)hipFree(dws);
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dxm);
    (void)hipFree(dwm);

    // triangular_solver: upper, non-unit, with a known solution (padded row strides)
    {
        using TS = linalg::triangular_solver<double, linalg::linalg_backend::ROCM_NATIVE>;
        const long long nt = 70, nr = 3, lda = nt + 2, ldb = nr + 1, ldx = nr + 3;
        std::vector<double> at(nt * lda, 0.0), bt(nt * ldb, -123.0), xt(nt * ldx, -123.0), want(nt * nr);
        for (long long i = 0; i < nt; ++i) {
            for (long long j = i; j < nt; ++j) at[i * lda + j] = i == j ? 2.0 + 0.02 * i : 0.01 * std::sin(double(i * nt + j + 1));
            for (long long j = 0; j < nr; ++j) want[i * nr + j] = std::sin(double(i * nr + j + 1));
        }
        for (long long i = 0; i < nt; ++i)
            for (long long j = 0; j < nr; ++j) {
                double s = 0.0;
                for (long long k = i; k < nt; ++k) s += at[i * lda + k] * want[k * nr + j];
                bt[i * ldb + j] = s;
            }
        double *d_a = nullptr, *d_b = nullptr, *d_x = nullptr;
        void* d_w = nullptr;
        const size_t wst = TS::workspace_size(nt, nr);
        HIP_OK(hipMalloc(&d_a, at.size() * sizeof(double)));
        HIP_OK(hipMalloc(&d_b, bt.size() * sizeof(double)));
        HIP_OK(hipMalloc(&d_x, xt.size() * sizeof(double)));
        HIP_OK(hipMalloc(&d_w, wst));
        HIP_OK(hipMemcpy(d_a, at.data(), at.size() * sizeof(double), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d_b, bt.data(), bt.size() * sizeof(double), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d_x, xt.data(), xt.size() * sizeof(double), hipMemcpyHostToDevice));
        TS::apply(true, false, false, nt, nr, d_a, lda, d_b, ldb, d_x, ldx, d_w, wst, nullptr);
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipMemcpy(xt.data(), d_x, xt.size() * sizeof(double), hipMemcpyDeviceToHost));
        double err = 0.0;
        bool pad = true;
        for (long long i = 0; i < nt; ++i) {
            for (long long j = 0; j < nr; ++j) err = std::fmax(err, std::fabs(xt[i * ldx + j] - want[i * nr + j]));
            for (long long j = nr; j < ldx; ++j) pad = pad && xt[i * ldx + j] == -123.0;
        }
        std::printf("triangular_solver n=%lld nrhs=%lld upper: max error %.2e\n", nt, nr, err);
        EXPECT(err <= 1e-13);
        EXPECT(pad);
        (void)hipFree(d_a);
        (void)hipFree(d_b);
        (void)hipFree(d_x);
        (void)hipFree(d_w);
    }

    // determinant and inversion of the reference's GPU case (TestGpuMatrixOperations.cxx:1466-1492):
    // A = diag(4, 2) gives det 8 and inverse diag(0.25, 0.5); a row swap flips the sign
    {
        using DT = linalg::determinant<double, linalg::linalg_backend::ROCM_NATIVE>;
        using IV = linalg::inversion<double, linalg::linalg_backend::ROCM_NATIVE>;
        const double diag[4] = {4.0, 0.0, 0.0, 2.0}, swapped[4] = {0.0, 2.0, 4.0, 0.0};
        double* d_m = nullptr;
        void* d_w = nullptr;
        const size_t wsd = std::max(DT::workspace_size(2, linear_solver_type::LU_LINEAR_SOLVER),
            IV::workspace_size(2, linear_solver_type::LU_LINEAR_SOLVER));
        HIP_OK(hipMalloc(&d_m, sizeof(diag)));
        HIP_OK(hipMalloc(&d_w, wsd));
        HIP_OK(hipMemcpy(d_m, diag, sizeof(diag), hipMemcpyHostToDevice));
        const double det = DT::apply(d_m, nullptr, 2, linear_solver_type::LU_LINEAR_SOLVER, d_w, wsd);
        HIP_OK(hipMemcpy(d_m, swapped, sizeof(swapped), hipMemcpyHostToDevice));
        const double det_sw = DT::apply(d_m, nullptr, 2, linear_solver_type::LU_LINEAR_SOLVER, d_w, wsd);
        HIP_OK(hipMemcpy(d_m, diag, sizeof(diag), hipMemcpyHostToDevice));
        IV::apply(d_m, nullptr, 2, linear_solver_type::LU_LINEAR_SOLVER, d_w, wsd);
        double inv[4] = {0, 0, 0, 0};
        HIP_OK(hipMemcpy(inv, d_m, sizeof(inv), hipMemcpyDeviceToHost));
        std::printf("determinant diag(4,2): %.17g, row-swapped: %.17g; inverse %g %g %g %g\n", det, det_sw, inv[0],
            inv[1], inv[2], inv[3]);
        EXPECT(det == 8.0);
        EXPECT(det_sw == -8.0);
        EXPECT(inv[0] == 0.25 && inv[1] == 0.0 && inv[2] == 0.0 && inv[3] == 0.5);
        // batched multiplication of two 2 x 2 pairs: [[1,2],[3,4]] [[5,6],[7,8]] and I * B
        const double ha[8] = {1, 2, 3, 4, 1, 0, 0, 1}, hb[8] = {5, 6, 7, 8, 2, 3, 4, 5};
        double hc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        double *d_a = nullptr, *d_b = nullptr, *d_c = nullptr;
        HIP_OK(hipMalloc(&d_a, sizeof(ha)));
        HIP_OK(hipMalloc(&d_b, sizeof(hb)));
        HIP_OK(hipMalloc(&d_c, sizeof(hc)));
        HIP_OK(hipMemcpy(d_a, ha, sizeof(ha), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d_b, hb, sizeof(hb), hipMemcpyHostToDevice));
        linalg::rocm_native::batched_matrix_multiplication(2, 2, d_a, d_b, d_c);
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipMemcpy(hc, d_c, sizeof(hc), hipMemcpyDeviceToHost));
        EXPECT(hc[0] == 19 && hc[1] == 22 && hc[2] == 43 && hc[3] == 50);
        EXPECT(hc[4] == 2 && hc[5] == 3 && hc[6] == 4 && hc[7] == 5);
        (void)hipFree(d_a);
        (void)hipFree(d_b);
        (void)hipFree(d_c);
        (void)hipFree(d_m);
        (void)hipFree(d_w);
    }
    {
        // cholesky_decomposition<ROCM_NATIVE>: A = C C^T with C lower triangular, 5 on the diagonal and
        // small positive entries below (the shape of the reference's known answer,
        // TestCholeskyDecomposition.cxx:170-210), gives back C (LOWER) and C^T (UPPER) with device
        // info 0; diag(4, -1) fails at order 2.
        using CH = linalg::cholesky_decomposition<double, linalg::linalg_backend::ROCM_NATIVE>;
        const long long nc = 70;
        std::vector<double> c(nc * nc, 0.0), a(nc * nc, 0.0);
        Normal u(77);
        for (long long i = 0; i < nc; ++i)
            for (long long j = 0; j <= i; ++j) c[i * nc + j] = (i == j) ? 5.0 : 0.5 * std::fabs(u()) / 3.0;
        for (long long i = 0; i < nc; ++i)
            for (long long j = 0; j < nc; ++j) {
                double acc = 0.0;
                for (long long p = 0; p < nc; ++p) acc += c[i * nc + p] * c[j * nc + p];
                a[i * nc + j] = acc;
            }
        for (int up = 0; up < 2; ++up) {
            const auto tri = up ? linalg::cholesky_decomposition_enum::UPPER_TRIANGULAR
                                : linalg::cholesky_decomposition_enum::LOWER_TRIANGULAR;
            const size_t wsb = CH::workspace_size(nc, tri);
            double* d_c = nullptr;
            int* d_info = nullptr;
            void* d_w = nullptr;
            HIP_OK(hipMalloc(&d_c, sizeof(double) * nc * nc));
            HIP_OK(hipMalloc(&d_info, sizeof(int)));
            HIP_OK(hipMalloc(&d_w, wsb));
            HIP_OK(hipMemcpy(d_c, a.data(), sizeof(double) * nc * nc, hipMemcpyHostToDevice));
            HIP_OK(hipMemset(d_info, 0x7f, sizeof(int)));
            EXPECT(CH::apply(d_c, nc, tri, d_w, wsb, d_info));
            HIP_OK(hipDeviceSynchronize());
            std::vector<double> f(nc * nc);
            int hinfo = -1;
            HIP_OK(hipMemcpy(f.data(), d_c, sizeof(double) * nc * nc, hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(&hinfo, d_info, sizeof(int), hipMemcpyDeviceToHost));
            double err = 0.0;
            for (long long i = 0; i < nc; ++i)
                for (long long j = 0; j < nc; ++j) {
                    const double want = up ? c[j * nc + i] : c[i * nc + j];
                    err = std::fmax(err, std::fabs(f[i * nc + j] - want));
                }
            std::printf("cholesky<ROCM_NATIVE> %s n=%lld: info %d, max|factor - C| %.2e\n", up ? "upper" : "lower", nc,
                hinfo, err);
            EXPECT(hinfo == 0);
            EXPECT(err <= 200 * 2.220446049250313e-16 * 5.0);
            const double bad[4] = {4.0, 0.0, 0.0, -1.0};
            HIP_OK(hipMemcpy(d_c, bad, sizeof(bad), hipMemcpyHostToDevice));
            EXPECT(CH::apply(d_c, 2, tri, d_w, wsb, d_info));
            HIP_OK(hipDeviceSynchronize());
            HIP_OK(hipMemcpy(&hinfo, d_info, sizeof(int), hipMemcpyDeviceToHost));
            EXPECT(hinfo == 2);
            (void)hipFree(d_c);
            (void)hipFree(d_info);
            (void)hipFree(d_w);
        }
    }
}

int main(int argc, char** argv)
{
    const bool no_gpu = argc > 1 && std::strcmp(argv[1], "--no-gpu") == 0;
    host_checks();
    if (!no_gpu) gpu_checks();
    if (g_fail == 0) std::printf("all checks passed (%s)\n", no_gpu ? "host only" : "host + GPU");
    return g_fail == 0 ? 0 : 1;
}
