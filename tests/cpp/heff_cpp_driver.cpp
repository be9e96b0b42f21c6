// Drives the C++ drop-in (include/dmrg_amd/matvec_mpos.hpp) the way the reference's callers do:
// PRIMME's matrixMatvec function pointer with the operator as context, MultAx(T*, T*), and the
// free matrix_vector_product / expectation_value. Reads inputs written by tests/test_gpu_cpp_adaptor.py,
// writes the results for comparison with the oracle.
//
// input : int64 header {cplx, d, chiL, chiR, DL, DR, nvec, ld}, then cx64 mpo, envL, envR, then x (ld x nvec, T)
// output: y_primme (ld x nvec, T), y_single (n, T), y_free (n, T), expval (2 doubles), num_mv (int64),
//         y_precond (ld x nvec, T; Jacobi with shifts 0.3, -0.7, 1.1, ...), y_shifted (n, T; H - 0.25 after
//         set_shift(0.25)), num_pc (int64)
#include "dmrg_amd/matvec_mpos.hpp"

static void *g_op = nullptr; // the preconditioner callback's operator

// Stand-in for PRIMME's parameter block: only the field the preconditioner callback reads.
struct primme_params {
    double *ShiftsForPreconditioner;
};

// The operator requirements of eig::solver::eigs_primme (solver_primme.cpp:297-517) and of the Jacobi callback
// (optimize_energy.cpp:34-42), exercised the way they use them.
enum class Preconditioner { NONE, JACOBI };
template <typename M>
static int solver_requirements(M &matrix, void *x, int ld, void *y, int nvec, double *shifts) {
    using Scalar = typename M::Scalar;
    static_assert(M::can_shift && M::can_precondition && M::can_shift_invert, "operator capabilities");
    static_assert(sizeof(Scalar) == 8 || sizeof(Scalar) == 16, "fp64 or cx64");
    if(matrix.isReadyFactorOp()) return 7;
    if(matrix.isReadyShift()) return 8;
    matrix.preconditioner = Preconditioner::JACOBI;
    matrix.set_jcbMaxBlockSize(1);
    primme_params pp{shifts};
    int           ldx = ld, ldy = ld, bs = nvec, err = 1;
    void (*pc)(void *, int *, void *, int *, int *, primme_params *, int *) = [](void *a, int *la, void *b, int *lb, int *n,
                                                                               primme_params *p, int *e) {
        static_cast<M *>(g_op)->MultPc(a, la, b, lb, n, p, e);
    };
    g_op = &matrix;
    pc(x, &ldx, y, &ldy, &bs, &pp, &err);
    if(err != 0) return 9;
    if(matrix.t_multOPv->get_time() != 0.0 || matrix.num_op != 0) return 10;
    return 0;
}

#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <vector>

using cx = std::complex<double>;

static void *g_matrix = nullptr; // stands in for primme->matrix

template <typename T>
static void primme_matvec_trampoline(void *x, int *ldx, void *y, int *ldy, int *bs, primme_params *pp, int *err) {
    static_cast<dmrg_amd::MatVecMPOS<T> *>(g_matrix)->MultAx(x, ldx, y, ldy, bs, pp, err);
}

template <typename T>
static int run(FILE *in, FILE *out, const int64_t h[8]) {
    const int64_t d = h[1], cL = h[2], cR = h[3], DL = h[4], DR = h[5], nvec = h[6], ld = h[7];
    const int64_t n = d * cL * cR;
    std::vector<cx> mpo(DL * DR * d * d), eL(cL * cL * DL), eR(cR * cR * DR);
    std::vector<T>  x(ld * nvec), y(ld * nvec, T(0)), y1(n), y2(n);
    if(fread(mpo.data(), sizeof(cx), mpo.size(), in) != mpo.size() || fread(eL.data(), sizeof(cx), eL.size(), in) != eL.size() ||
       fread(eR.data(), sizeof(cx), eR.size(), in) != eR.size() || fread(x.data(), sizeof(T), x.size(), in) != x.size())
        return 2;
    dmrg_amd::MatVecMPOS<T> op({mpo.data()}, {{DL, DR, d, d}}, eL.data(), {cL, cL, DL}, eR.data(), {cR, cR, DR});
    if(op.rows() != n) return 3;
    g_matrix = &op;
    void (*matvec)(void *, int *, void *, int *, int *, primme_params *, int *) = primme_matvec_trampoline<T>;
    int ldx = static_cast<int>(ld), ldy = static_cast<int>(ld), bs = static_cast<int>(nvec), err = 1;
    matvec(x.data(), &ldx, y.data(), &ldy, &bs, nullptr, &err);
    if(err != 0) return 4;
    op.MultAx(x.data(), y1.data());
    // Free function on the same first vector, with the operator's own (possibly real-cast) tensors.
    std::vector<T> mpo_t(mpo.size()), eL_t(eL.size()), eR_t(eR.size());
    if constexpr(std::is_same_v<T, double>) {
        for(size_t i = 0; i < mpo.size(); ++i) mpo_t[i] = mpo[i].real();
        for(size_t i = 0; i < eL.size(); ++i) eL_t[i] = eL[i].real();
        for(size_t i = 0; i < eR.size(); ++i) eR_t[i] = eR[i].real();
    } else {
        mpo_t = mpo, eL_t = eL, eR_t = eR;
    }
    dmrg_amd::contraction::matrix_vector_product<T>(y2.data(), x.data(), {d, cL, cR}, mpo_t.data(), {DL, DR, d, d}, eL_t.data(),
                                                    {cL, cL, DL}, eR_t.data(), {cR, cR, DR});
    const T ev = dmrg_amd::contraction::expectation_value<T>(x.data(), {d, cL, cR}, mpo_t.data(), {DL, DR, d, d}, eL_t.data(),
                                                             {cL, cL, DL}, eR_t.data(), {cR, cR, DR});
    const double  evr[2] = {std::real(cx(ev)), std::imag(cx(ev))};
    const int64_t nmv    = op.num_mv;
    fwrite(y.data(), sizeof(T), y.size(), out);
    fwrite(y1.data(), sizeof(T), y1.size(), out);
    fwrite(y2.data(), sizeof(T), y2.size(), out);
    fwrite(evr, sizeof(double), 2, out);
    fwrite(&nmv, sizeof(int64_t), 1, out);
    std::vector<T>      ypc(ld * nvec, T(0)), ysh(n);
    std::vector<double> shifts(nvec);
    for(int64_t i = 0; i < nvec; ++i) shifts[i] = (i % 3 == 0) ? 0.3 : (i % 3 == 1 ? -0.7 : 1.1);
    if(const int rc = solver_requirements(op, x.data(), static_cast<int>(ld), ypc.data(), static_cast<int>(nvec), shifts.data()))
        return rc;
    op.set_shift(cx(0.25, 0.0));
    if(!op.isReadyShift()) return 11;
    op.MultAx(x.data(), ysh.data());
    if(op.t_multAx->get_time() <= 0.0) return 12;
    const int64_t npc = op.num_pc;
    fwrite(ypc.data(), sizeof(T), ypc.size(), out);
    fwrite(ysh.data(), sizeof(T), ysh.size(), out);
    fwrite(&npc, sizeof(int64_t), 1, out);
    // MultBx: without a B operator the output is left as it is; with B built on the same blocks it equals MultAx.
    {
        std::vector<T> yb(n, T(7));
        op.MultBx(x.data(), yb.data());
        for(const T &v : yb)
            if(v != T(7)) return 13;
        dmrg_amd::MatVecMPOS<T> ab({mpo.data()}, {{DL, DR, d, d}}, eL.data(), {cL, cL, DL}, eR.data(), {cR, cR, DR}, eL.data(),
                                   {cL, cL, DL}, eR.data(), {cR, cR, DR});
        std::vector<T> ya(n), yb2(n);
        ab.MultAx(x.data(), ya.data());
        ab.MultBx(x.data(), yb2.data());
        double dmax = 0, amax = 0;
        for(int64_t i = 0; i < n; ++i) dmax = std::max(dmax, std::abs(ya[i] - yb2[i])), amax = std::max(amax, std::abs(ya[i]));
        if(dmax > 1e-13 * amax) return 14;
        if(ab.num_mv != 2) return 15;
    }
    // Error path: a dimension mismatch surfaces as dmrg_amd::error with DMRG_ERR_DIMS.
    try {
        dmrg_amd::contraction::matrix_vector_product<T>(y2.data(), x.data(), {d, cL + 1, cR}, mpo_t.data(), {DL, DR, d, d},
                                                        eL_t.data(), {cL, cL, DL}, eR_t.data(), {cR, cR, DR});
        return 5;
    } catch(const dmrg_amd::error &e) {
        if(e.status != DMRG_ERR_DIMS) return 6;
    }
    return 0;
}

int main(int argc, char **argv) {
    if(argc != 3) return 1;
    FILE *in = fopen(argv[1], "rb"), *out = fopen(argv[2], "wb");
    if(!in || !out) return 1;
    int64_t h[8];
    if(fread(h, sizeof(int64_t), 8, in) != 8) return 2;
    const int rc = h[0] ? run<cx>(in, out, h) : run<double>(in, out, h);
    fclose(in);
    fclose(out);
    if(rc) fprintf(stderr, "heff_cpp_driver failed with code %d\n", rc);
    return rc;
}
