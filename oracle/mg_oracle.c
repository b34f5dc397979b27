/*
 * CPU oracle (test infrastructure only): FP64 restatement of the reference's
 * matchgate kernels, brickwall sweeps, target assembly, Riemannian projection,
 * Hessian-vector product and trust region.  Only tests/, the smoke check and
 * bench.py's cpu_baseline leg call it; the product library does not.
 *
 * Each function names the reference file:line whose algorithm it restates.
 */
#include "mg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define C(g, k) ((g)[(k)])      /* center entry k = 0..3 */
#define K(g, k) ((g)[4 + (k)])  /* corner entry k = 0..3 */

/* amplitude indices of the quadruple t for the sorted wire pair (i < j):
 * x=|00>, y=|01>, z=|10>, w=|11> (first label = wire i, qubit 0 = MSB) */
static inline void quad_index(long t, int i, int j, int L, long* x, long* y, long* z, long* w)
{
    const int pi = L - 1 - i, pj = L - 1 - j; /* pi > pj */
    long a = t;
    a = ((a >> pj) << (pj + 1)) | (a & ((1L << pj) - 1));
    a = ((a >> pi) << (pi + 1)) | (a & ((1L << pi) - 1));
    *x = a;
    *y = a | (1L << pj);
    *z = a | (1L << pi);
    *w = a | (1L << pi) | (1L << pj);
}

/* reversed wires: center block [c3,c2,c1,c0], corner unchanged (reference src/gate.c:588-604) */
static inline void sort_gate(const ocplx* g, int* i, int* j, ocplx* gs, int* flipped)
{
    memcpy(gs, g, 8 * sizeof(ocplx));
    *flipped = 0;
    if (*j < *i) {
        gs[0] = g[3]; gs[1] = g[2]; gs[2] = g[1]; gs[3] = g[0];
        int t = *i; *i = *j; *j = t;
        *flipped = 1;
    }
}

static inline void unflip_grad(ocplx* dg)
{
    ocplx t = dg[0]; dg[0] = dg[3]; dg[3] = t;
    t = dg[1]; dg[1] = dg[2]; dg[2] = t;
}

/* reference src/gate.c:581-652 (apply_matchgate) */
void orc_apply_gate(const ocplx* g0, int i, int j, int L, const ocplx* psi, ocplx* out)
{
    ocplx g[8]; int fl;
    sort_gate(g0, &i, &j, g, &fl);
    const long nq = 1L << (L - 2);
    for (long t = 0; t < nq; t++) {
        long ix, iy, iz, iw;
        quad_index(t, i, j, L, &ix, &iy, &iz, &iw);
        const ocplx x = psi[ix], y = psi[iy], z = psi[iz], w = psi[iw];
        out[ix] = K(g, 0) * x + K(g, 1) * w;
        out[iy] = C(g, 0) * y + C(g, 1) * z;
        out[iz] = C(g, 2) * y + C(g, 3) * z;
        out[iw] = K(g, 2) * x + K(g, 3) * w;
    }
}

/* reference src/gate.c:659-772 (apply_matchgate_backward): dpsi = V^T dout, dg = dout (x) psi */
void orc_apply_gate_backward(const ocplx* g0, int i, int j, int L, const ocplx* psi,
                             const ocplx* dout, ocplx* dpsi, ocplx* dg)
{
    ocplx g[8]; int fl;
    sort_gate(g0, &i, &j, g, &fl);
    for (int k = 0; k < 8; k++) dg[k] = 0;
    const long nq = 1L << (L - 2);
    for (long t = 0; t < nq; t++) {
        long ix, iy, iz, iw;
        quad_index(t, i, j, L, &ix, &iy, &iz, &iw);
        const ocplx x = psi[ix], y = psi[iy], z = psi[iz], w = psi[iw];
        const ocplx dx = dout[ix], dy = dout[iy], dz = dout[iz], dw = dout[iw];
        K(dg, 0) += dx * x; K(dg, 1) += dx * w;
        C(dg, 0) += dy * y; C(dg, 1) += dy * z;
        C(dg, 2) += dz * y; C(dg, 3) += dz * z;
        K(dg, 2) += dw * x; K(dg, 3) += dw * w;
        dpsi[ix] = K(g, 0) * dx + K(g, 2) * dw;
        dpsi[iy] = C(g, 0) * dy + C(g, 2) * dz;
        dpsi[iz] = C(g, 1) * dy + C(g, 3) * dz;
        dpsi[iw] = K(g, 1) * dx + K(g, 3) * dw;
    }
    if (fl) unflip_grad(dg);
}

/* reference src/gate.c:779-851 (apply_matchgate_to_array), interleaved data[idx*nstates + k] */
void orc_apply_gate_to_array(const ocplx* g0, int i, int j, int L, int ns, const ocplx* psi, ocplx* out)
{
    ocplx g[8]; int fl;
    sort_gate(g0, &i, &j, g, &fl);
    const long nq = 1L << (L - 2);
    for (long t = 0; t < nq; t++) {
        long ix, iy, iz, iw;
        quad_index(t, i, j, L, &ix, &iy, &iz, &iw);
        for (int k = 0; k < ns; k++) {
            const ocplx x = psi[ix * ns + k], y = psi[iy * ns + k], z = psi[iz * ns + k], w = psi[iw * ns + k];
            out[ix * ns + k] = K(g, 0) * x + K(g, 1) * w;
            out[iy * ns + k] = C(g, 0) * y + C(g, 1) * z;
            out[iz * ns + k] = C(g, 2) * y + C(g, 3) * z;
            out[iw * ns + k] = K(g, 2) * x + K(g, 3) * w;
        }
    }
}

/* reference src/gate.c:858-973 (apply_matchgate_backward_array): dgs[k] = dout (x) psi_k */
void orc_apply_gate_backward_array(const ocplx* g0, int i, int j, int L, int ns, const ocplx* psi,
                                   const ocplx* dout, ocplx* dgs)
{
    ocplx g[8]; int fl;
    sort_gate(g0, &i, &j, g, &fl);
    memset(dgs, 0, (size_t)ns * 8 * sizeof(ocplx));
    const long nq = 1L << (L - 2);
    for (long t = 0; t < nq; t++) {
        long ix, iy, iz, iw;
        quad_index(t, i, j, L, &ix, &iy, &iz, &iw);
        const ocplx dx = dout[ix], dy = dout[iy], dz = dout[iz], dw = dout[iw];
        for (int k = 0; k < ns; k++) {
            ocplx* dg = dgs + 8 * k;
            const ocplx x = psi[ix * ns + k], y = psi[iy * ns + k], z = psi[iz * ns + k], w = psi[iw * ns + k];
            K(dg, 0) += dx * x; K(dg, 1) += dx * w;
            C(dg, 0) += dy * y; C(dg, 1) += dy * z;
            C(dg, 2) += dz * y; C(dg, 3) += dz * z;
            K(dg, 2) += dw * x; K(dg, 3) += dw * w;
        }
    }
    if (fl)
        for (int k = 0; k < ns; k++) unflip_grad(dgs + 8 * k);
}

/* reference src/gate.c:982-1057 (apply_matchgate_placeholder): the 8 states d(V psi)/dV[e] */
void orc_apply_gate_placeholder(int i, int j, int L, const ocplx* psi, ocplx* out)
{
    const long S = 1L << L;
    memset(out, 0, (size_t)S * 8 * sizeof(ocplx));
    int fl = 0;
    if (j < i) { int t = i; i = j; j = t; fl = 1; }
    const long nq = 1L << (L - 2);
    for (long t = 0; t < nq; t++) {
        long ix, iy, iz, iw;
        quad_index(t, i, j, L, &ix, &iy, &iz, &iw);
        const ocplx x = psi[ix], y = psi[iy], z = psi[iz], w = psi[iw];
        /* sorted-basis center entries e = 0..3 map to the gate's entry (fl ? 3-e : e) */
        out[iy * 8 + (fl ? 3 : 0)] = y;
        out[iy * 8 + (fl ? 2 : 1)] = z;
        out[iz * 8 + (fl ? 1 : 2)] = y;
        out[iz * 8 + (fl ? 0 : 3)] = z;
        out[ix * 8 + 4] = x;
        out[ix * 8 + 5] = w;
        out[iw * 8 + 6] = x;
        out[iw * 8 + 7] = w;
    }
}

/* ---------------- gate programs ---------------- */

/* reference src/matchgate_target.c:308-324 (matchgate_brickwall_to_sequential) */
void orc_brickwall_program(int L, int nlayers, const int* perms, int* param, int* wires)
{
    const int per = L / 2;
    for (int i = 0; i < nlayers; i++)
        for (int j = 0; j < per; j++) {
            param[i * per + j] = i;
            wires[2 * (i * per + j)] = perms[i * L + 2 * j];
            wires[2 * (i * per + j) + 1] = perms[i * L + 2 * j + 1];
        }
}

/* reference src/quantum_circuit.c:13-35 */
void orc_apply_program(const ocplx* gates, int G, const int* param, const int* wires, int L,
                       const ocplx* psi, ocplx* out)
{
    const long S = 1L << L;
    ocplx* a = malloc(S * sizeof(ocplx));
    ocplx* b = malloc(S * sizeof(ocplx));
    memcpy(a, psi, S * sizeof(ocplx));
    for (int g = 0; g < G; g++) {
        orc_apply_gate(gates + 8 * param[g], wires[2 * g], wires[2 * g + 1], L, a, b);
        ocplx* t = a; a = b; b = t;
    }
    memcpy(out, a, S * sizeof(ocplx));
    free(a); free(b);
}

/* reference src/quantum_circuit.c:90-112, src/matchgate_brickwall.c:52-78 */
void orc_program_forward(const ocplx* gates, int G, const int* param, const int* wires, int L,
                         const ocplx* psi, ocplx* cache, ocplx* out)
{
    const long S = 1L << L;
    memcpy(cache, psi, S * sizeof(ocplx));
    for (int g = 0; g < G; g++)
        orc_apply_gate(gates + 8 * param[g], wires[2 * g], wires[2 * g + 1], L, cache + g * S,
                       g + 1 < G ? cache + (g + 1) * S : out);
}

/* reference src/quantum_circuit.c:115-140, src/matchgate_brickwall.c:87-127 */
void orc_program_backward(const ocplx* gates, int G, const int* param, const int* wires, int L,
                          const ocplx* cache, const ocplx* dout, ocplx* dpsi, ocplx* dgi)
{
    const long S = 1L << L;
    ocplx* a = malloc(S * sizeof(ocplx));
    ocplx* b = malloc(S * sizeof(ocplx));
    memcpy(a, dout, S * sizeof(ocplx));
    for (int g = G - 1; g >= 0; g--) {
        orc_apply_gate_backward(gates + 8 * param[g], wires[2 * g], wires[2 * g + 1], L, cache + g * S, a, b, dgi + 8 * g);
        ocplx* t = a; a = b; b = t;
    }
    memcpy(dpsi, a, S * sizeof(ocplx));
    free(a); free(b);
}

/* reference src/matchgate_brickwall.c:137-260 (matchgate_brickwall_unitary_backward_hessian),
 * generalised to gate programs: A[(P(p),x),(P(q),y)] += h_pq[x][y] over instance pairs p < q,
 * hess = A + A^T.  For the tied brickwall this is the reference's diagonal-block
 * symmetrisation (:233-243) plus mirror (:244-252) operation for operation. */
void orc_program_backward_hessian(const ocplx* gates, int P, int G, const int* param, const int* wires, int L,
                                  const ocplx* cache, const ocplx* dout, ocplx* dpsi, ocplx* dgrad, ocplx* hess)
{
    const long S = 1L << L;
    const int m = 8 * P;
    /* gradient with the per-instance output gradients kept (grad_cache[k] = lambda_{k+1}) */
    ocplx* gc = malloc((size_t)G * S * sizeof(ocplx));
    memcpy(gc + (size_t)(G - 1) * S, dout, S * sizeof(ocplx));
    memset(dgrad, 0, (size_t)m * sizeof(ocplx));
    for (int g = G - 1; g >= 0; g--) {
        ocplx dg[8];
        orc_apply_gate_backward(gates + 8 * param[g], wires[2 * g], wires[2 * g + 1], L, cache + g * S,
                                gc + (size_t)g * S, g > 0 ? gc + (size_t)(g - 1) * S : dpsi, dg);
        for (int k = 0; k < 8; k++) dgrad[8 * param[g] + k] += dg[k];
    }
    ocplx* A = calloc((size_t)m * m, sizeof(ocplx));
    ocplx* t0 = malloc((size_t)S * 8 * sizeof(ocplx));
    ocplx* t1 = malloc((size_t)S * 8 * sizeof(ocplx));
    ocplx h[64];
    for (int p = 0; p < G; p++) {
        orc_apply_gate_placeholder(wires[2 * p], wires[2 * p + 1], L, cache + (size_t)p * S, t0);
        ocplx *cur = t0, *nxt = t1;
        for (int q = p + 1; q < G; q++) {
            orc_apply_gate_backward_array(gates + 8 * param[q], wires[2 * q], wires[2 * q + 1], L, 8, cur,
                                          gc + (size_t)q * S, h);
            const int i = param[p], jj = param[q];
            for (int x = 0; x < 8; x++)
                for (int y = 0; y < 8; y++) A[(size_t)(8 * i + x) * m + 8 * jj + y] += h[8 * x + y];
            if (q < G - 1) {
                orc_apply_gate_to_array(gates + 8 * param[q], wires[2 * q], wires[2 * q + 1], L, 8, cur, nxt);
                ocplx* t = cur; cur = nxt; nxt = t;
            }
        }
    }
    for (int a = 0; a < m; a++)
        for (int b = 0; b < m; b++) hess[(size_t)a * m + b] = A[(size_t)a * m + b] + A[(size_t)b * m + a];
    free(A); free(t0); free(t1); free(gc);
}

/* reference src/matchgate_brickwall.c:266-306 (matchgate_ti_brickwall_unitary_backward): the
 * gradient of each layer from its last pair only (cache state k = (i+1) L/2 - 1, wires
 * perms[i][L-2], perms[i][L-1]), scaled by L/2; the backpropagated state passes through those
 * pairs only.  The reference alternates two buffers by the parity of k, which chains the
 * layers in this order when L/2 is odd; for even L/2 it reads a buffer not written by the
 * previous layer, and this restatement keeps the chain. */
void orc_ti_brickwall_backward(const ocplx* vlist, int nlayers, const int* perms, int L, const ocplx* cache,
                               const ocplx* dout, ocplx* dpsi, ocplx* dgrad)
{
    const long S = 1L << L;
    const int per = L / 2;
    ocplx* cur = malloc(S * sizeof(ocplx));
    ocplx* nxt = malloc(S * sizeof(ocplx));
    memcpy(cur, dout, S * sizeof(ocplx));
    for (int i = nlayers - 1; i >= 0; i--) {
        const int k = (i + 1) * per - 1;
        ocplx dg[8];
        orc_apply_gate_backward(vlist + 8 * i, perms[(size_t)i * L + L - 2], perms[(size_t)i * L + L - 1], L,
                                cache + (size_t)k * S, cur, nxt, dg);
        for (int e = 0; e < 8; e++) dgrad[8 * i + e] = per * dg[e];
        ocplx* t = cur; cur = nxt; nxt = t;
    }
    memcpy(dpsi, cur, S * sizeof(ocplx));
    free(cur); free(nxt);
}

/* reference src/matchgate_brickwall.c:308-435 (matchgate_ti_brickwall_unitary_backward_hessian) */
void orc_ti_brickwall_backward_hessian(const ocplx* vlist, int nlayers, const int* perms, int L,
                                       const ocplx* cache, const ocplx* dout, ocplx* dpsi, ocplx* dgrad, ocplx* hess)
{
    const long S = 1L << L;
    const int per = L / 2, G = nlayers * per, m = 8 * nlayers;
    int* param = malloc(G * sizeof(int));
    int* wires = malloc(2 * G * sizeof(int));
    orc_brickwall_program(L, nlayers, perms, param, wires);
    ocplx* gc = malloc((size_t)G * S * sizeof(ocplx));
    memcpy(gc + (size_t)(G - 1) * S, dout, S * sizeof(ocplx));
    memset(dgrad, 0, (size_t)m * sizeof(ocplx));
    for (int g = G - 1; g >= 0; g--) {
        ocplx dg[8];
        orc_apply_gate_backward(vlist + 8 * param[g], wires[2 * g], wires[2 * g + 1], L, cache + g * S,
                                gc + (size_t)g * S, g > 0 ? gc + (size_t)(g - 1) * S : dpsi, dg);
        for (int k = 0; k < 8; k++) dgrad[8 * param[g] + k] += dg[k];
    }
    ocplx* A = calloc((size_t)m * m, sizeof(ocplx));
    ocplx* t0 = malloc((size_t)S * 8 * sizeof(ocplx));
    ocplx* t1 = malloc((size_t)S * 8 * sizeof(ocplx));
    ocplx h[64];
    for (int i = 0; i < nlayers; i++) {
        const int p = i * per; /* placeholder only at pair r = 0 */
        orc_apply_gate_placeholder(wires[2 * p], wires[2 * p + 1], L, cache + (size_t)p * S, t0);
        ocplx *cur = t0, *nxt = t1;
        for (int q = p + 1; q < G; q++) {
            const int j = param[q], s = q - j * per;
            double mult = per;
            if (i == j) mult -= s; /* o = 1, 2, ... over s = 1, 2, ... */
            orc_apply_gate_backward_array(vlist + 8 * j, wires[2 * q], wires[2 * q + 1], L, 8, cur, gc + (size_t)q * S, h);
            for (int x = 0; x < 8; x++)
                for (int y = 0; y < 8; y++) A[(size_t)(8 * i + x) * m + 8 * j + y] += mult * h[8 * x + y];
            if (q < G - 1) {
                orc_apply_gate_to_array(vlist + 8 * j, wires[2 * q], wires[2 * q + 1], L, 8, cur, nxt);
                ocplx* t = cur; cur = nxt; nxt = t;
            }
        }
    }
    for (int a = 0; a < m; a++)
        for (int b = 0; b < m; b++) hess[(size_t)a * m + b] = A[(size_t)a * m + b] + A[(size_t)b * m + a];
    free(A); free(t0); free(t1); free(gc); free(param); free(wires);
}

/* ---------------- target assembly over basis columns ---------------- */

static int nthreads_or_default(int nthreads)
{
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

static inline int thread_id(void)
{
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

/* seed Upsi = -conj(U e_b) (reference src/matchgate_target.c:441-455) */
static inline void seed_column(const ocplx* U, long S, long b, ocplx* s)
{
    for (long a = 0; a < S; a++) s[a] = -conj(U[a * S + b]);
}

/* reference src/matchgate_target.c:20-80 and :83-203 (OpenMP over columns, per-thread partials);
 * the partials are merged in thread order so that the result is reproducible */
void orc_target(const ocplx* U, const ocplx* gates, int G, const int* param, const int* wires, int L,
                int nthreads, ocplx* f)
{
    const long S = 1L << L;
    const int nt = nthreads_or_default(nthreads);
    ocplx* part = calloc(nt, sizeof(ocplx));
#pragma omp parallel num_threads(nt)
    {
        ocplx* psi = malloc(S * sizeof(ocplx));
        ocplx* s = malloc(S * sizeof(ocplx));
        ocplx* out = malloc(S * sizeof(ocplx));
        ocplx acc = 0;
#pragma omp for schedule(static)
        for (long b = 0; b < S; b++) {
            memset(psi, 0, S * sizeof(ocplx));
            psi[b] = 1;
            seed_column(U, S, b, s);
            orc_apply_program(gates, G, param, wires, L, psi, out);
            for (long a = 0; a < S; a++) acc += s[a] * out[a];
        }
        part[thread_id()] = acc;
        free(psi); free(s); free(out);
    }
    ocplx tot = 0;
    for (int t = 0; t < nt; t++) tot += part[t];
    *f = tot;
    free(part);
}

/* reference src/matchgate_target.c:212-302 and :359-388 (per-instance gradients summed per parameter) */
void orc_target_gradient(const ocplx* U, const ocplx* gates, int P, int G, const int* param, const int* wires, int L,
                         int nthreads, ocplx* f, ocplx* dgrad)
{
    const long S = 1L << L;
    const int nt = nthreads_or_default(nthreads);
    ocplx* part = calloc((size_t)nt * (1 + 8 * P), sizeof(ocplx));
#pragma omp parallel num_threads(nt)
    {
        ocplx* psi = malloc(S * sizeof(ocplx));
        ocplx* s = malloc(S * sizeof(ocplx));
        ocplx* out = malloc(S * sizeof(ocplx));
        ocplx* cache = malloc((size_t)G * S * sizeof(ocplx));
        ocplx* dgi = malloc((size_t)G * 8 * sizeof(ocplx));
        ocplx* my = part + (size_t)thread_id() * (1 + 8 * P);
#pragma omp for schedule(static)
        for (long b = 0; b < S; b++) {
            memset(psi, 0, S * sizeof(ocplx));
            psi[b] = 1;
            seed_column(U, S, b, s);
            orc_program_forward(gates, G, param, wires, L, psi, cache, out);
            for (long a = 0; a < S; a++) my[0] += s[a] * out[a];
            orc_program_backward(gates, G, param, wires, L, cache, s, psi, dgi);
            for (int g = 0; g < G; g++)
                for (int k = 0; k < 8; k++) my[1 + 8 * param[g] + k] += dgi[8 * g + k];
        }
        free(psi); free(s); free(out); free(cache); free(dgi);
    }
    ocplx tf = 0;
    for (int k = 0; k < 8 * P; k++) dgrad[k] = 0;
    for (int t = 0; t < nt; t++) {
        tf += part[(size_t)t * (1 + 8 * P)];
        for (int k = 0; k < 8 * P; k++) dgrad[k] += part[(size_t)t * (1 + 8 * P) + 1 + k];
    }
    *f = tf;
    free(part);
}

/* reference src/matchgate_target.c:397-513 (serial) and :626-869 (OpenMP) */
void orc_target_gradient_hessian(const ocplx* U, const ocplx* gates, int P, int G, const int* param, const int* wires,
                                 int L, int ti, int nthreads, ocplx* f, ocplx* dgrad, ocplx* hess)
{
    const long S = 1L << L;
    const int m = 8 * P;
    const int nt = nthreads_or_default(nthreads);
    const size_t stride = 1 + (size_t)m + (size_t)m * m;
    ocplx* part = calloc((size_t)nt * stride, sizeof(ocplx));
    int* perms = NULL;
    if (ti) { /* recover the tied layer permutations from the program */
        const int per = L / 2;
        perms = malloc((size_t)P * L * sizeof(int));
        for (int g = 0; g < G; g++) {
            perms[param[g] * L + 2 * (g % per)] = wires[2 * g];
            perms[param[g] * L + 2 * (g % per) + 1] = wires[2 * g + 1];
        }
    }
#pragma omp parallel num_threads(nt)
    {
        ocplx* psi = malloc(S * sizeof(ocplx));
        ocplx* s = malloc(S * sizeof(ocplx));
        ocplx* out = malloc(S * sizeof(ocplx));
        ocplx* cache = malloc((size_t)G * S * sizeof(ocplx));
        ocplx* dg = malloc((size_t)m * sizeof(ocplx));
        ocplx* hu = malloc((size_t)m * m * sizeof(ocplx));
        ocplx* my = part + (size_t)thread_id() * stride;
#pragma omp for schedule(static)
        for (long b = 0; b < S; b++) {
            memset(psi, 0, S * sizeof(ocplx));
            psi[b] = 1;
            seed_column(U, S, b, s);
            orc_program_forward(gates, G, param, wires, L, psi, cache, out);
            for (long a = 0; a < S; a++) my[0] += s[a] * out[a];
            if (ti)
                orc_ti_brickwall_backward_hessian(gates, P, perms, L, cache, s, psi, dg, hu);
            else
                orc_program_backward_hessian(gates, P, G, param, wires, L, cache, s, psi, dg, hu);
            for (int k = 0; k < m; k++) my[1 + k] += dg[k];
            for (size_t k = 0; k < (size_t)m * m; k++) my[1 + m + k] += hu[k];
        }
        free(psi); free(s); free(out); free(cache); free(dg); free(hu);
    }
    ocplx tf = 0;
    memset(dgrad, 0, (size_t)m * sizeof(ocplx));
    memset(hess, 0, (size_t)m * m * sizeof(ocplx));
    for (int t = 0; t < nt; t++) {
        const ocplx* p = part + (size_t)t * stride;
        tf += p[0];
        for (int k = 0; k < m; k++) dgrad[k] += p[1 + k];
        for (size_t k = 0; k < (size_t)m * m; k++) hess[k] += p[1 + m + k];
    }
    *f = tf;
    free(part); free(perms);
}

/* Forward-over-reverse directional second derivative (no reference function: the
 * reference declares an HVP optimiser, src/trust_region.h:84-86, whose body is
 * commented out).  Per column b with seed s = -conj(U e_b):
 *   forward  psi_{g+1} = V_g psi_g,  dpsi_{g+1} = V_g dpsi_g + Z_g psi_g
 *   f_b = s . psi_G
 *   reverse  lam_G = s, dlam_G = 0
 *            grad_g += lam_{g+1} (x) psi_g
 *            dhz_g  += dlam_{g+1} (x) psi_g + lam_{g+1} (x) dpsi_g
 *            lam_g = V_g^T lam_{g+1},  dlam_g = V_g^T dlam_{g+1} + Z_g^T lam_{g+1}
 * The outer products and transposed applications are the reference's backward
 * kernel (src/gate.c:659-772). */
/* seed s = -conj(T e_b) for a target program T (gate k on (twires[2k], twires[2k+1])), the
 * per-column ufunc of the matrix-free drivers (examples/hubbard1d/hubbard1d_opt.c:30-39) */
static void seed_program(const ocplx* tgates, int Gt, const int* twires, int L, long b, ocplx* s, ocplx* tmp)
{
    const long S = 1L << L;
    memset(tmp, 0, S * sizeof(ocplx));
    tmp[b] = 1;
    ocplx* in = tmp;
    ocplx* out = s;
    for (int g = 0; g < Gt; g++) {
        orc_apply_gate(tgates + 8 * g, twires[2 * g], twires[2 * g + 1], L, in, out);
        ocplx* t = in; in = out; out = t;
    }
    for (long a = 0; a < S; a++) s[a] = -conj(in[a]);
}

static void target_gradient_hvp_core(const ocplx* U, const ocplx* tgates, int Gt, const int* twires,
                                     const ocplx* gates, const ocplx* Z, int P, int G, const int* param,
                                     const int* wires, int L, long b0, long b1, int nthreads,
                                     ocplx* f, ocplx* dgrad, ocplx* dhz);

void orc_target_gradient_hvp(const ocplx* U, const ocplx* gates, const ocplx* Z, int P, int G, const int* param,
                             const int* wires, int L, long b0, long b1, int nthreads,
                             ocplx* f, ocplx* dgrad, ocplx* dhz)
{
    target_gradient_hvp_core(U, NULL, 0, NULL, gates, Z, P, G, param, wires, L, b0, b1, nthreads, f, dgrad, dhz);
}

void orc_target_gradient_hvp_program(const ocplx* tgates, int Gt, const int* twires, const ocplx* gates, const ocplx* Z,
                                     int P, int G, const int* param, const int* wires, int L, long b0, long b1,
                                     int nthreads, ocplx* f, ocplx* dgrad, ocplx* dhz)
{
    target_gradient_hvp_core(NULL, tgates, Gt, twires, gates, Z, P, G, param, wires, L, b0, b1, nthreads, f, dgrad,
                             dhz);
}

static void target_gradient_hvp_core(const ocplx* U, const ocplx* tgates, int Gt, const int* twires,
                                     const ocplx* gates, const ocplx* Z, int P, int G, const int* param,
                                     const int* wires, int L, long b0, long b1, int nthreads,
                                     ocplx* f, ocplx* dgrad, ocplx* dhz)
{
    const long S = 1L << L;
    const int m = 8 * P;
    const int nt = nthreads_or_default(nthreads);
    const size_t stride = 1 + 2 * (size_t)m;
    ocplx* part = calloc((size_t)nt * stride, sizeof(ocplx));
#pragma omp parallel num_threads(nt)
    {
        ocplx* s = malloc(S * sizeof(ocplx));
        ocplx* cp = malloc((size_t)(G + 1) * S * sizeof(ocplx));  /* psi_0 .. psi_G */
        ocplx* cd = malloc((size_t)(G + 1) * S * sizeof(ocplx));  /* dpsi_0 .. dpsi_G */
        ocplx* lam = malloc(S * sizeof(ocplx));
        ocplx* dlam = malloc(S * sizeof(ocplx));
        ocplx* t0 = malloc(S * sizeof(ocplx));
        ocplx* t1 = malloc(S * sizeof(ocplx));
        ocplx* t2 = malloc(S * sizeof(ocplx));
        ocplx* my = part + (size_t)thread_id() * stride;
#pragma omp for schedule(static)
        for (long b = b0; b < b1; b++) {
            if (U) seed_column(U, S, b, s);
            else seed_program(tgates, Gt, twires, L, b, s, t0);
            memset(cp, 0, S * sizeof(ocplx));
            memset(cd, 0, S * sizeof(ocplx));
            cp[b] = 1;
            for (int g = 0; g < G; g++) {
                const ocplx* V = gates + 8 * param[g];
                const ocplx* Zg = Z + 8 * param[g];
                const int i = wires[2 * g], j = wires[2 * g + 1];
                orc_apply_gate(V, i, j, L, cp + (size_t)g * S, cp + (size_t)(g + 1) * S);
                orc_apply_gate(V, i, j, L, cd + (size_t)g * S, cd + (size_t)(g + 1) * S);
                orc_apply_gate(Zg, i, j, L, cp + (size_t)g * S, t0);
                for (long a = 0; a < S; a++) cd[(size_t)(g + 1) * S + a] += t0[a];
            }
            ocplx fb = 0;
            for (long a = 0; a < S; a++) fb += s[a] * cp[(size_t)G * S + a];
            my[0] += fb;
            memcpy(lam, s, S * sizeof(ocplx));
            memset(dlam, 0, S * sizeof(ocplx));
            for (int g = G - 1; g >= 0; g--) {
                const ocplx* V = gates + 8 * param[g];
                const ocplx* Zg = Z + 8 * param[g];
                const int i = wires[2 * g], j = wires[2 * g + 1];
                ocplx d1[8], d2[8], d3[8];
                /* lam (x) psi_g and V^T lam */
                orc_apply_gate_backward(V, i, j, L, cp + (size_t)g * S, lam, t0, d1);
                /* dlam (x) psi_g and V^T dlam */
                orc_apply_gate_backward(V, i, j, L, cp + (size_t)g * S, dlam, t1, d2);
                /* lam (x) dpsi_g and Z^T lam (the outer product does not depend on the gate) */
                orc_apply_gate_backward(Zg, i, j, L, cd + (size_t)g * S, lam, t2, d3);
                for (long a = 0; a < S; a++) dlam[a] = t1[a] + t2[a];
                memcpy(lam, t0, S * sizeof(ocplx));
                for (int k = 0; k < 8; k++) {
                    my[1 + 8 * param[g] + k] += d1[k];
                    my[1 + m + 8 * param[g] + k] += d2[k] + d3[k];
                }
            }
        }
        free(s); free(cp); free(cd); free(lam); free(dlam); free(t0); free(t1); free(t2);
    }
    ocplx tf = 0;
    memset(dgrad, 0, (size_t)m * sizeof(ocplx));
    memset(dhz, 0, (size_t)m * sizeof(ocplx));
    for (int t = 0; t < nt; t++) {
        const ocplx* p = part + (size_t)t * stride;
        tf += p[0];
        for (int k = 0; k < m; k++) { dgrad[k] += p[1 + k]; dhz[k] += p[1 + m + k]; }
    }
    *f = tf;
    free(part);
}

/* ---------------- matchgate algebra (reference src/matchgate.c) ---------------- */

static void mul2(const ocplx* a, const ocplx* b, ocplx* c)
{
    c[0] = a[0] * b[0] + a[1] * b[2];
    c[1] = a[0] * b[1] + a[1] * b[3];
    c[2] = a[2] * b[0] + a[3] * b[2];
    c[3] = a[2] * b[1] + a[3] * b[3];
}

/* reference src/matchgate.c:356-378 */
void orc_multiply(const ocplx* a, const ocplx* b, ocplx* c)
{
    ocplx t[8];
    mul2(a, b, t);
    mul2(a + 4, b + 4, t + 4);
    memcpy(c, t, sizeof(t));
}

/* reference src/matchgate.c:182-194 */
void orc_adjoint(const ocplx* a, ocplx* ah)
{
    ocplx t[8];
    for (int blk = 0; blk < 2; blk++) {
        const ocplx* s = a + 4 * blk;
        t[4 * blk + 0] = conj(s[0]); t[4 * blk + 1] = conj(s[2]);
        t[4 * blk + 2] = conj(s[1]); t[4 * blk + 3] = conj(s[3]);
    }
    memcpy(ah, t, sizeof(t));
}

/* reference src/matchgate.c:93-104 */
void orc_conjugate(ocplx* a)
{
    for (int k = 0; k < 8; k++) a[k] = conj(a[k]);
}

/* reference src/matchgate.c:199-217: Hermitian part of each block */
void orc_symm(const ocplx* w, ocplx* z)
{
    ocplx t[8];
    for (int blk = 0; blk < 2; blk++) {
        const ocplx* s = w + 4 * blk;
        ocplx* d = t + 4 * blk;
        d[0] = creal(s[0]);
        d[3] = creal(s[3]);
        d[1] = 0.5 * (s[1] + conj(s[2]));
        d[2] = conj(d[1]);
    }
    memcpy(z, t, sizeof(t));
}

/* reference src/matchgate.c:222-246: skew-Hermitian part of each block */
void orc_antisymm(const ocplx* w, ocplx* z)
{
    ocplx t[8];
    for (int blk = 0; blk < 2; blk++) {
        const ocplx* s = w + 4 * blk;
        ocplx* d = t + 4 * blk;
        d[0] = I * cimag(s[0]);
        d[3] = I * cimag(s[3]);
        d[1] = 0.5 * (s[1] - conj(s[2]));
        d[2] = -conj(d[1]);
    }
    memcpy(z, t, sizeof(t));
}

/* reference src/matchgate.c:251-272 */
void orc_real_to_antisymm(const double* r, ocplx* w)
{
    for (int blk = 0; blk < 2; blk++) {
        const double* q = r + 4 * blk;
        ocplx* d = w + 4 * blk;
        d[0] = I * q[0];
        d[3] = I * q[3];
        d[1] = 0.5 * (q[1] - q[2]) + 0.5 * I * (q[1] + q[2]);
        d[2] = -conj(d[1]);
    }
}

/* reference src/matchgate.c:277-290 */
void orc_antisymm_to_real(const ocplx* w, double* r)
{
    for (int k = 0; k < 8; k++) r[k] = creal(w[k]) + cimag(w[k]);
}

/* reference src/matchgate.c:296-301 */
void orc_real_to_tangent(const double* r, const ocplx* v, ocplx* z)
{
    ocplx a[8];
    orc_real_to_antisymm(r, a);
    orc_multiply(v, a, z);
}

/* reference src/matchgate.c:309-316 */
void orc_tangent_to_real(const ocplx* v, const ocplx* z, double* r)
{
    ocplx w[8], t[8];
    orc_adjoint(v, w);
    orc_multiply(w, z, t);
    orc_antisymm(t, w);
    orc_antisymm_to_real(w, r);
}

/* reference src/matchgate.c:323-335: p = z - u symm(u^dagger z) */
void orc_project_tangent(const ocplx* u, const ocplx* z, ocplx* p)
{
    ocplx v[8], w[8];
    orc_adjoint(u, v);
    orc_multiply(v, z, w);
    orc_symm(w, v);
    orc_multiply(u, v, w);
    for (int k = 0; k < 8; k++) p[k] = z[k] - w[k];
}

/* reference src/matchgate.c:381-394: (a b c + c b a) / 2 */
void orc_symmetric_triple(const ocplx* a, const ocplx* b, const ocplx* c, ocplx* ret)
{
    ocplx u[8], v[8], r[8];
    orc_multiply(a, b, u);
    orc_multiply(u, c, r);
    orc_multiply(c, b, u);
    orc_multiply(u, a, v);
    for (int k = 0; k < 8; k++) ret[k] = 0.5 * (r[k] + v[k]);
}

/* reference src/matchgate.c:414-489: Gauss-Jordan elimination with partial pivoting */
int orc_inverse_mat2x2(const ocplx* a, ocplx* ainv)
{
    ocplx m[4] = { a[0], a[1], a[2], a[3] };
    ocplx v[4] = { 1, 0, 0, 1 };
    for (int k = 0; k < 2; k++) {
        int piv = k;
        double p = cabs(m[2 * k + k]);
        if (k == 0 && cabs(m[2]) > p) { piv = 1; p = cabs(m[2]); }
        if (p == 0) return -1;
        if (piv != k) {
            for (int c = 0; c < 2; c++) {
                ocplx t = m[2 * k + c]; m[2 * k + c] = m[2 * piv + c]; m[2 * piv + c] = t;
                t = v[2 * k + c]; v[2 * k + c] = v[2 * piv + c]; v[2 * piv + c] = t;
            }
        }
        for (int r = 0; r < 2; r++) {
            if (r == k) continue;
            const ocplx s = m[2 * r + k] / m[2 * k + k];
            for (int c = k + 1; c < 2; c++) m[2 * r + c] -= s * m[2 * k + c];
            m[2 * r + k] = 0;
            for (int c = 0; c < 2; c++) v[2 * r + c] -= s * v[2 * k + c];
        }
        for (int c = k + 1; c < 2; c++) m[2 * k + c] /= m[2 * k + k];
        for (int c = 0; c < 2; c++) v[2 * k + c] /= m[2 * k + k];
        m[2 * k + k] = 1;
    }
    memcpy(ainv, v, sizeof(v));
    return 0;
}

static double udist(int n, const ocplx* x, const ocplx* y)
{
    double d = 0;
    for (int k = 0; k < n; k++) d = fmax(d, cabs(x[k] - y[k]));
    return d;
}

/* reference src/matchgate.c:517-543: Newton iteration u <- (u + u^{-dagger}) / 2 (Higham 1986) */
void orc_polar(const ocplx* a, ocplx* u)
{
    memcpy(u, a, 8 * sizeof(ocplx));
    for (int k = 0; k < 14; k++) {
        ocplx v[8], w[8];
        orc_inverse_mat2x2(u, v);
        orc_inverse_mat2x2(u + 4, v + 4);
        orc_adjoint(v, w);
        for (int e = 0; e < 8; e++) u[e] = 0.5 * (u[e] + w[e]);
        if (k >= 4 && udist(4, u, w) < 1e-14 && udist(4, u + 4, w + 4) < 1e-14) break;
    }
}

/* reference src/matchgate.c:548-562: polar factor of u (I + A(eta)) */
void orc_retract(const ocplx* u, const double* eta, ocplx* v)
{
    ocplx z[8], w[8];
    orc_real_to_antisymm(eta, z);
    z[0] += 1; z[3] += 1; z[4] += 1; z[7] += 1;
    orc_multiply(u, z, w);
    orc_polar(w, v);
}

/* ---------------- Riemannian projection ---------------- */

/* reference src/matchgate_target.c:903-908 */
void orc_riemannian_gradient(const ocplx* vlist, int P, const ocplx* dgrad, double* grad_vec)
{
    for (int i = 0; i < P; i++) {
        ocplx g[8];
        memcpy(g, dgrad + 8 * i, sizeof(g));
        orc_conjugate(g);
        orc_tangent_to_real(vlist + 8 * i, g, grad_vec + 8 * i);
    }
}

/* reference src/matchgate_target.c:877-973 */
void orc_gradient_vector_hessian_matrix(const ocplx* vlist, int P, const ocplx* dgrad, const ocplx* hess,
                                        double* grad_vec, double* H)
{
    const int m = 8 * P;
    orc_riemannian_gradient(vlist, P, dgrad, grad_vec);
    for (int i = 0; i < P; i++)
        for (int j = i; j < P; j++)
            for (int k = 0; k < 8; k++) {
                double r[8] = { 0 };
                r[k] = 1;
                ocplx Z[8], G[8] = { 0 };
                orc_real_to_tangent(r, vlist + 8 * j, Z);
                for (int x = 0; x < 8; x++)
                    for (int y = 0; y < 8; y++) G[x] += hess[(size_t)(8 * i + x) * m + 8 * j + y] * Z[y];
                orc_conjugate(G);
                if (i == j) {
                    ocplx Gp[8], gh[8], gc[8], T[8];
                    orc_project_tangent(vlist + 8 * i, G, Gp);
                    memcpy(gc, dgrad + 8 * i, sizeof(gc));
                    orc_conjugate(gc);
                    orc_adjoint(gc, gh);
                    orc_symmetric_triple(Z, gh, vlist + 8 * i, T);
                    for (int e = 0; e < 8; e++) G[e] = Gp[e] - T[e];
                }
                orc_tangent_to_real(vlist + 8 * i, G, r);
                for (int x = 0; x < 8; x++) H[(size_t)(8 * i + x) * m + 8 * j + k] = r[x];
            }
    for (int a = 0; a < m; a++)
        for (int b = 0; b < a; b++) H[(size_t)a * m + b] = H[(size_t)b * m + a];
}

/* tangent directions Z_i = V_i A(v_i) (reference src/matchgate.c:296-301) */
void orc_real_to_directions(const ocplx* vlist, int P, const double* v, ocplx* Z)
{
    for (int i = 0; i < P; i++) orc_real_to_tangent(v + 8 * i, vlist + 8 * i, Z + 8 * i);
}

/* Riemannian HVP from the complex directional derivative dhz = D^2 f[Z]: the same
 * projection the reference applies per Hessian column (src/matchgate_target.c:910-961),
 * applied to a whole direction at once; equals H v when every V_i is unitary */
void orc_riemannian_hvp(const ocplx* vlist, int P, const ocplx* dgrad, const ocplx* Z, const ocplx* dhz, double* hvp)
{
    for (int i = 0; i < P; i++) {
        ocplx G[8], Gp[8], gc[8], gh[8], T[8];
        memcpy(G, dhz + 8 * i, sizeof(G));
        orc_conjugate(G);
        orc_project_tangent(vlist + 8 * i, G, Gp);
        memcpy(gc, dgrad + 8 * i, sizeof(gc));
        orc_conjugate(gc);
        orc_adjoint(gc, gh);
        orc_symmetric_triple(Z + 8 * i, gh, vlist + 8 * i, T);
        for (int e = 0; e < 8; e++) G[e] = Gp[e] - T[e];
        orc_tangent_to_real(vlist + 8 * i, G, hvp + 8 * i);
    }
}

/* ---------------- trust region (reference src/trust_region.c) ---------------- */

static double dot(int n, const double* a, const double* b)
{
    double s = 0;
    for (int k = 0; k < n; k++) s += a[k] * b[k];
    return s;
}

/* reference src/trust_region.c:19-45 */
static void solve_quadratic(double p, double q, double x[2])
{
    if (p == 0) { double a = sqrt(-q); x[0] = -a; x[1] = a; return; }
    const double sg = copysign(1.0, p);
    const double a = -(p + sg * sqrt(p * p - q));
    const double b = q / a;
    if (a <= b) { x[0] = a; x[1] = b; } else { x[0] = b; x[1] = a; }
}

/* reference src/trust_region.c:54-72 */
static double move_to_boundary(const double* b, const double* d, int n, double radius)
{
    const double dsq = dot(n, d, d);
    if (dsq == 0) return 0;
    const double p = dot(n, b, d) / dsq;
    const double q = (dot(n, b, b) - radius * radius) / dsq;
    double t[2];
    solve_quadratic(p, q, t);
    return t[1];
}

typedef void (*orc_hvp_fn)(void* ctx, const double* d, double* Hd);

/* reference src/trust_region.c:88-159 (hmat) and :164-229 (hvp): Steihaug-Toint tCG */
static int tcg(const double* grad, orc_hvp_fn hv, void* ctx, int n, double radius,
               int maxiter, double abstol, double reltol, double* z, int* nhvp)
{
    double* r = malloc(n * sizeof(double));
    double* d = malloc(n * sizeof(double));
    double* Hd = malloc(n * sizeof(double));
    memcpy(r, grad, n * sizeof(double));
    double rsq = dot(n, r, r);
    const double stoptol = fmax(abstol, reltol * sqrt(rsq));
    memset(z, 0, n * sizeof(double));
    for (int k = 0; k < n; k++) d[k] = -r[k];
    int boundary = 0;
    for (int it = 0; it < maxiter; it++) {
        hv(ctx, d, Hd);
        if (nhvp) (*nhvp)++;
        const double dHd = dot(n, d, Hd);
        const double t = move_to_boundary(z, d, n, radius);
        const double alpha = rsq / dHd;
        if (dHd <= 0 || alpha > t) {
            for (int k = 0; k < n; k++) z[k] += t * d[k];
            boundary = 1;
            break;
        }
        for (int k = 0; k < n; k++) r[k] += alpha * Hd[k];
        for (int k = 0; k < n; k++) z[k] += alpha * d[k];
        const double rsq_next = dot(n, r, r);
        if (sqrt(rsq_next) <= stoptol) break;
        const double beta = rsq_next / rsq;
        for (int k = 0; k < n; k++) d[k] = beta * d[k] - r[k];
        rsq = rsq_next;
    }
    free(r); free(d); free(Hd);
    return boundary;
}

struct hmat_ctx { const double* H; int n; };

static void hmat_mv(void* c, const double* d, double* Hd)
{
    const struct hmat_ctx* h = c;
    for (int a = 0; a < h->n; a++) Hd[a] = dot(h->n, h->H + (size_t)a * h->n, d);
}

int orc_truncated_cg_hmat(const double* grad, const double* hess, int n, double radius,
                          int maxiter, double abstol, double reltol, double* z)
{
    struct hmat_ctx c = { hess, n };
    return tcg(grad, hmat_mv, &c, n, radius, maxiter, abstol, reltol, z, NULL);
}

struct hvp_ctx {
    const ocplx* U; const ocplx* V; const ocplx* dgrad;
    int P, G, L, nthreads; const int* param; const int* wires;
};

static void program_hvp(void* c, const double* v, double* hv)
{
    const struct hvp_ctx* h = c;
    const int m = 8 * h->P;
    ocplx* Z = malloc(m * sizeof(ocplx));
    ocplx* dg = malloc(m * sizeof(ocplx));
    ocplx* dhz = malloc(m * sizeof(ocplx));
    ocplx f;
    orc_real_to_directions(h->V, h->P, v, Z);
    orc_target_gradient_hvp(h->U, h->V, Z, h->P, h->G, h->param, h->wires, h->L, 0, 1L << h->L, h->nthreads, &f, dg, dhz);
    orc_riemannian_hvp(h->V, h->P, h->dgrad, Z, dhz, hv);
    free(Z); free(dg); free(dhz);
}

/* reference src/trust_region.c:235-306 (hmat); the HVP mode implements the optimiser
 * declared at src/trust_region.h:84-86 (its body is commented out, :319-381) */
void orc_rtr_optimize(const ocplx* U, const ocplx* vstart, int P, int G, const int* param, const int* wires, int L,
                      int mode, int niter, int tcg_maxiter, double rho_trust, double radius_init, double maxradius,
                      double abstol, double reltol, int nthreads, double* f_iter, ocplx* vopt, int* hvp_count)
{
    const int n = 8 * P;
    ocplx* x = malloc((size_t)n * sizeof(ocplx));
    ocplx* xn = malloc((size_t)n * sizeof(ocplx));
    memcpy(x, vstart, (size_t)n * sizeof(ocplx));
    double* grad = malloc(n * sizeof(double));
    double* H = mode == 0 ? malloc((size_t)n * n * sizeof(double)) : NULL;
    double* eta = malloc(n * sizeof(double));
    double* Heta = malloc(n * sizeof(double));
    ocplx* dgrad = malloc(n * sizeof(ocplx));
    ocplx* hess = mode == 0 ? malloc((size_t)n * n * sizeof(ocplx)) : NULL;
    double radius = radius_init;
    if (hvp_count) *hvp_count = 0;
    for (int k = 0; k < niter; k++) {
        ocplx f;
        int bnd;
        if (mode == 0) {
            orc_target_gradient_hessian(U, x, P, G, param, wires, L, 0, nthreads, &f, dgrad, hess);
            orc_gradient_vector_hessian_matrix(x, P, dgrad, hess, grad, H);
            bnd = orc_truncated_cg_hmat(grad, H, n, radius, tcg_maxiter, abstol, reltol, eta);
            struct hmat_ctx c = { H, n };
            hmat_mv(&c, eta, Heta);
        } else {
            orc_target_gradient(U, x, P, G, param, wires, L, nthreads, &f, dgrad);
            orc_riemannian_gradient(x, P, dgrad, grad);
            struct hvp_ctx c = { U, x, dgrad, P, G, L, nthreads, param, wires };
            bnd = tcg(grad, program_hvp, &c, n, radius, tcg_maxiter, abstol, reltol, eta, hvp_count);
            program_hvp(&c, eta, Heta);
            if (hvp_count) (*hvp_count)++;
        }
        for (int i = 0; i < P; i++) orc_retract(x + 8 * i, eta + 8 * i, xn + 8 * i);
        const double fx = creal(f);
        f_iter[k] = fx;
        ocplx fn;
        orc_target(U, xn, G, param, wires, L, nthreads, &fn);
        const double rho = (creal(fn) - fx) / (dot(n, grad, eta) + 0.5 * dot(n, eta, Heta));
        if (rho < 0.25) radius *= 0.25;
        else if (rho > 0.75 && bnd) radius = fmin(2 * radius, maxradius);
        if (rho > rho_trust) { ocplx* t = x; x = xn; xn = t; }
    }
    memcpy(vopt, x, (size_t)n * sizeof(ocplx));
    ocplx f;
    orc_target(U, x, G, param, wires, L, nthreads, &f);
    f_iter[niter] = creal(f);
    free(x); free(xn); free(grad); free(H); free(eta); free(Heta); free(dgrad); free(hess);
}
