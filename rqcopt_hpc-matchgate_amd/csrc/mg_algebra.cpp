// Host FP64 2x2-block algebra of matchgates, statevector allocation and vector
// distances (C-ABI of include/matchgate.h, include/statevector.h, include/util.h).
// Each function states the reference function whose contract it keeps.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/matchgate.h"
#include "../../include/statevector.h"
#include "../../include/util.h"

namespace {

using zd = std::complex<double>;

inline zd* blk(struct matchgate* a, int b) { return b ? a->corner_block.data : a->center_block.data; }
inline const zd* blk(const struct matchgate* a, int b) { return b ? a->corner_block.data : a->center_block.data; }

// c = a b for row-major 2x2 blocks
inline void mul22(const zd* a, const zd* b, zd* c)
{
    const zd c0 = a[0] * b[0] + a[1] * b[2];
    const zd c1 = a[0] * b[1] + a[1] * b[3];
    const zd c2 = a[2] * b[0] + a[3] * b[2];
    const zd c3 = a[2] * b[1] + a[3] * b[3];
    c[0] = c0; c[1] = c1; c[2] = c2; c[3] = c3;
}

}  // namespace

extern "C" {

// ---- matchgate algebra (reference src/matchgate.c) ----

void compare_matchgate_arrays(const struct matchgate* m, const struct matchgate* n, int d)
{
    std::printf("\n");
    for (int i = 0; i < d; i++) {
        std::printf("\nMatchgate %d:\n", i);
        for (int b = 0; b < 2; b++)
            for (int k = 0; k < 4; k++) {
                const zd x = blk(&m[i], b)[k], y = blk(&n[i], b)[k];
                std::printf("%s %d: %lf + %lfi           %lf + %lfi\n", b ? "Corner" : "Center", k + 4 * b, x.real(),
                            x.imag(), y.real(), y.imag());
            }
    }
}

void identity_mat2x2(struct mat2x2* a)
{
    a->data[0] = 1.0; a->data[1] = 0.0; a->data[2] = 0.0; a->data[3] = 1.0;
}

void identity_matchgate(struct matchgate* a)
{
    identity_mat2x2(&a->center_block);
    identity_mat2x2(&a->corner_block);
}

void scale_matchgate(struct matchgate* a, const double x)
{
    for (int b = 0; b < 2; b++)
        for (int k = 0; k < 4; k++) blk(a, b)[k] *= x;
}

void conjugate_matchgate(struct matchgate* a)
{
    for (int b = 0; b < 2; b++)
        for (int k = 0; k < 4; k++) blk(a, b)[k] = std::conj(blk(a, b)[k]);
}

void add_matchgate(struct matchgate* a, const struct matchgate* b)
{
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < 4; k++) blk(a, q)[k] += blk(b, q)[k];
}

void sub_matchgate(struct matchgate* a, const struct matchgate* b)
{
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < 4; k++) blk(a, q)[k] -= blk(b, q)[k];
}

void add_matchgates(const struct matchgate* a, const struct matchgate* b, struct matchgate* c)
{
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < 4; k++) blk(c, q)[k] = blk(a, q)[k] + blk(b, q)[k];
}

void sub_matchgates(const struct matchgate* a, const struct matchgate* b, struct matchgate* c)
{
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < 4; k++) blk(c, q)[k] = blk(a, q)[k] - blk(b, q)[k];
}

void transpose_matchgate(const struct matchgate* a, struct matchgate* at)
{
    for (int q = 0; q < 2; q++) {
        const zd* s = blk(a, q);
        zd* d = blk(at, q);
        const zd t1 = s[1], t2 = s[2];
        d[0] = s[0]; d[3] = s[3]; d[1] = t2; d[2] = t1;
    }
}

void adjoint_matchgate(const struct matchgate* a, struct matchgate* ah)
{
    for (int q = 0; q < 2; q++) {
        const zd* s = blk(a, q);
        zd* d = blk(ah, q);
        const zd t1 = std::conj(s[1]), t2 = std::conj(s[2]);
        d[0] = std::conj(s[0]); d[3] = std::conj(s[3]); d[1] = t2; d[2] = t1;
    }
}

// Hermitian part of each block (reference src/matchgate.c:199-217)
void symm_matchgate(const struct matchgate* w, struct matchgate* z)
{
    for (int q = 0; q < 2; q++) {
        const zd* s = blk(w, q);
        zd* d = blk(z, q);
        const zd off = 0.5 * (s[1] + std::conj(s[2]));
        const double d0 = s[0].real(), d3 = s[3].real();
        d[0] = d0; d[3] = d3; d[1] = off; d[2] = std::conj(off);
    }
}

// skew-Hermitian part of each block (reference src/matchgate.c:222-246)
void antisymm_matchgate(const struct matchgate* w, struct matchgate* z)
{
    for (int q = 0; q < 2; q++) {
        const zd* s = blk(w, q);
        zd* d = blk(z, q);
        const zd off = 0.5 * (s[1] - std::conj(s[2]));
        const zd d0(0.0, s[0].imag()), d3(0.0, s[3].imag());
        d[0] = d0; d[3] = d3; d[1] = off; d[2] = -std::conj(off);
    }
}

// isometry R^8 -> skew-Hermitian blocks (reference src/matchgate.c:251-272)
void real_to_antisymm_matchgate(const double* r, struct matchgate* w)
{
    for (int q = 0; q < 2; q++) {
        const double* x = r + 4 * q;
        zd* d = blk(w, q);
        d[0] = zd(0.0, x[0]);
        d[3] = zd(0.0, x[3]);
        d[1] = zd(0.5 * (x[1] - x[2]), 0.5 * (x[1] + x[2]));
        d[2] = -std::conj(d[1]);
    }
}

// inverse isometry (reference src/matchgate.c:277-290): r = Re + Im entrywise
void antisymm_matchgate_to_real(const struct matchgate* w, double* r)
{
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < 4; k++) r[4 * q + k] = blk(w, q)[k].real() + blk(w, q)[k].imag();
}

void multiply_mat2x2(const struct mat2x2* a, const struct mat2x2* b, struct mat2x2* c)
{
    mul22(a->data, b->data, c->data);
}

void multiply_matchgates(const struct matchgate* a, const struct matchgate* b, struct matchgate* c)
{
    mul22(a->center_block.data, b->center_block.data, c->center_block.data);
    mul22(a->corner_block.data, b->corner_block.data, c->corner_block.data);
}

// z = v A(r) (reference src/matchgate.c:296-301)
void real_to_tangent_matchgate(const double* r, const struct matchgate* v, struct matchgate* z)
{
    struct matchgate a;
    real_to_antisymm_matchgate(r, &a);
    multiply_matchgates(v, &a, z);
}

// r = coordinates of antisymm(v^dagger z) (reference src/matchgate.c:309-316)
void tangent_matchgate_to_real(const struct matchgate* v, const struct matchgate* z, double* r)
{
    struct matchgate vh, t, a;
    adjoint_matchgate(v, &vh);
    multiply_matchgates(&vh, z, &t);
    antisymm_matchgate(&t, &a);
    antisymm_matchgate_to_real(&a, r);
}

// p = z - u symm(u^dagger z) (reference src/matchgate.c:323-335)
void project_tangent_matchgate(const struct matchgate* u, const struct matchgate* z, struct matchgate* p)
{
    struct matchgate uh, w, s, us;
    adjoint_matchgate(u, &uh);
    multiply_matchgates(&uh, z, &w);
    symm_matchgate(&w, &s);
    multiply_matchgates(u, &s, &us);
    sub_matchgates(z, &us, p);
}

// (a b c + c b a) / 2 (reference src/matchgate.c:381-394)
void symmetric_triple_matchgate_product(const struct matchgate* a, const struct matchgate* b, const struct matchgate* c,
                                        struct matchgate* ret)
{
    struct matchgate u, v, w;
    multiply_matchgates(a, b, &u);
    multiply_matchgates(&u, c, &v);
    multiply_matchgates(c, b, &u);
    multiply_matchgates(&u, a, &w);
    add_matchgates(&v, &w, ret);
    scale_matchgate(ret, 0.5);
}

// Gauss-Jordan elimination with partial pivoting; -1 if singular (reference src/matchgate.c:414-489)
int inverse_mat2x2(const struct mat2x2* a, struct mat2x2* ainv)
{
    zd m[4] = {a->data[0], a->data[1], a->data[2], a->data[3]};
    zd v[4] = {1.0, 0.0, 0.0, 1.0};
    for (int k = 0; k < 2; k++) {
        int piv = k;
        double p = std::abs(m[2 * k + k]);
        if (k == 0 && std::abs(m[2]) > p) {
            piv = 1;
            p = std::abs(m[2]);
        }
        if (p == 0.0) return -1;
        if (p < 1e-12) std::fprintf(stderr, "warning: encountered an almost singular matrix in 'inverse_matrix', p = %g\n", p);
        if (piv != k)
            for (int c = 0; c < 2; c++) {
                std::swap(m[2 * k + c], m[2 * piv + c]);
                std::swap(v[2 * k + c], v[2 * piv + c]);
            }
        for (int r = 0; r < 2; r++) {
            if (r == k) continue;
            const zd s = m[2 * r + k] / m[2 * k + k];
            for (int c = k + 1; c < 2; c++) m[2 * r + c] -= s * m[2 * k + c];
            m[2 * r + k] = 0.0;
            for (int c = 0; c < 2; c++) v[2 * r + c] -= s * v[2 * k + c];
        }
        for (int c = k + 1; c < 2; c++) m[2 * k + c] /= m[2 * k + k];
        for (int c = 0; c < 2; c++) v[2 * k + c] /= m[2 * k + k];
        m[2 * k + k] = 1.0;
    }
    for (int c = 0; c < 4; c++) ainv->data[c] = v[c];
    return 0;
}

int inverse_matchgate(const struct matchgate* a, struct matchgate* ainv)
{
    const int r0 = inverse_mat2x2(&a->center_block, &ainv->center_block);
    const int r1 = inverse_mat2x2(&a->corner_block, &ainv->corner_block);
    return (r0 < 0 || r1 < 0) ? -1 : 0;
}

// unitary polar factor by the Newton iteration u <- (u + u^{-dagger}) / 2
// (Higham 1986; reference src/matchgate.c:517-543: at most 14 steps, stop after
// step 5 once both blocks moved less than 1e-14 in the max norm)
void polar_matchgate_factor(const struct matchgate* a, struct matchgate* u)
{
    *u = *a;
    for (int k = 0; k < 14; k++) {
        struct matchgate inv, w;
        inverse_matchgate(u, &inv);
        adjoint_matchgate(&inv, &w);
        add_matchgate(u, &w);
        scale_matchgate(u, 0.5);
        if (k >= 4 && uniform_distance(4, u->center_block.data, w.center_block.data) < 1e-14 &&
            uniform_distance(4, u->corner_block.data, w.corner_block.data) < 1e-14)
            break;
    }
}

// retraction: polar factor of u (I + A(eta)) (reference src/matchgate.c:548-562)
void retract_matchgate(const struct matchgate* u, const double* eta, struct matchgate* v)
{
    struct matchgate z, w;
    real_to_antisymm_matchgate(eta, &z);
    z.center_block.data[0] += 1.0;
    z.center_block.data[3] += 1.0;
    z.corner_block.data[0] += 1.0;
    z.corner_block.data[3] += 1.0;
    multiply_matchgates(u, &z, &w);
    polar_matchgate_factor(&w, v);
}

// ---- distances (reference src/util.c:8-49) ----

double uniform_distance_real(long n, const double* x, const double* y)
{
    double d = 0.0;
    for (long i = 0; i < n; i++) d = std::fmax(d, std::fabs(x[i] - y[i]));
    return d;
}

double uniform_distance(long n, const numeric* x, const numeric* y)
{
    double d = 0.0;
    for (long i = 0; i < n; i++) d = std::fmax(d, std::abs(x[i] - y[i]));
    return d;
}

double relative_distance(long n, const numeric* x, const numeric* y, double eps)
{
    double d = 0.0;
    for (long i = 0; i < n; i++) d = std::fmax(d, std::abs(x[i] - y[i]) / std::fmax(std::abs(x[i]) + std::abs(y[i]), eps));
    return d;
}

// ---- statevectors (reference src/statevector.c) ----

int allocate_statevector(int nqubits, struct statevector* psi)
{
    psi->nqubits = nqubits;
    const size_t bytes = ((size_t)1 << nqubits) * sizeof(numeric);
    psi->data = (numeric*)aligned_alloc(MEM_DATA_ALIGN, std::max<size_t>(bytes, MEM_DATA_ALIGN));
    if (!psi->data) {
        std::fprintf(stderr, "allocating statevector memory failed\n");
        return -1;
    }
    return 0;
}

void free_statevector(struct statevector* psi)
{
    free(psi->data);
    psi->data = nullptr;
    psi->nqubits = 0;
}

// psi_trans[idx'] = psi[idx] with idx' the index whose bit for qubit perm[k] is bit k of idx
// (reference src/statevector.c:49-74)
void transpose_statevector(const struct statevector* psi, const int* perm, struct statevector* psi_trans)
{
    const int L = psi->nqubits;
    const size_t S = (size_t)1 << L;
    for (size_t a = 0; a < S; a++) {
        size_t b = 0;
        for (int k = 0; k < L; k++) {
            const size_t bit = (a >> (L - 1 - perm[k])) & 1u;
            b |= bit << (L - 1 - k);
        }
        psi_trans->data[b] = psi->data[a];
    }
}

void add_statevectors(const struct statevector* psi1, const struct statevector* psi2, struct statevector* psi_sum)
{
    const size_t S = (size_t)1 << psi1->nqubits;
    for (size_t a = 0; a < S; a++) psi_sum->data[a] = psi1->data[a] + psi2->data[a];
}

int allocate_statevector_array(int nqubits, int nstates, struct statevector_array* psi_list)
{
    psi_list->nqubits = nqubits;
    psi_list->nstates = nstates;
    const size_t bytes = ((size_t)1 << nqubits) * nstates * sizeof(numeric);
    psi_list->data = (numeric*)aligned_alloc(MEM_DATA_ALIGN, std::max<size_t>(bytes, MEM_DATA_ALIGN));
    if (!psi_list->data) {
        std::fprintf(stderr, "allocating statevector array memory failed\n");
        return -1;
    }
    return 0;
}

void free_statevector_array(struct statevector_array* psi_list)
{
    free(psi_list->data);
    psi_list->data = nullptr;
    psi_list->nqubits = 0;
    psi_list->nstates = 0;
}

}  // extern "C"
