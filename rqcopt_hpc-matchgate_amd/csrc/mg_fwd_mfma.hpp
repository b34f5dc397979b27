// Forward passes of sweeps on the matrix cores: a 12-qubit LDS tile (a whole L=12 column, or one
// tile of a larger one) swept through 4-qubit windows, each window's product of gates applied as
// two 8 x 8 parity blocks by v_mfma_f32_16x16x4f32 (FP32: struct FWin and the fragments made by
// forward_windows) or v_mfma_f64_16x16x4f64 (FP64: struct DWin, forward_windows64; mg_mat_mfma.hip).
// The tile's 16 column blocks (16 columns of the 256 that share the window's complement) are
// shared by the NW waves; each wave reads all its operands of a window before writing them back in
// place, then one barrier per window.
#pragma once
#include "mg_mat_mfma.hpp"

namespace mg {

typedef float v4f __attribute__((ext_vector_type(4)));

// psi <- M psi over the windows [w0, w1); fragments [window][8][lane]
template <int NW>
__device__ inline void fwd_windows(const FWin* fwin, const float* fwaf, int w0, int w1, cpx<float>* t_psi, int lane,
                                   int wv)
{
    constexpr int CB = 16 / NW;  // column blocks per wave
    const int n0 = lane >> 4, col = lane & 15;
    for (int w = w0; w < w1; w++) {
        const FWin* fw = fwin + w;
        float am[8];
#pragma unroll
        for (int f = 0; f < 8; f++) am[f] = fwaf[((size_t)w * 8 + f) * 64 + lane];
        const uint32_t oj = (uint32_t)fw->offj[col];
        uint32_t base[CB];
        cpx<float> x[CB][2][2];
#pragma unroll
        for (int c = 0; c < CB; c++) {
            base[c] = oj ^ (uint32_t)fw->offcb[wv * CB + c];
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int h = 0; h < 2; h++) x[c][p][h] = t_psi[base[c] ^ (uint32_t)fw->offp[p][4 * h + n0]];
        }
        v4f ap[CB][2];
#pragma unroll
        for (int c = 0; c < CB; c++)
#pragma unroll
            for (int p = 0; p < 2; p++) ap[c][p] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
#pragma unroll
            for (int c = 0; c < CB; c++)
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const cpx<float>& xv = x[c][p][kk >> 1];
                    ap[c][p] = __builtin_amdgcn_mfma_f32_16x16x4f32(am[4 * p + kk], (kk & 1) ? xv.im : xv.re, ap[c][p],
                                                                    0, 0, 0);
                }
        // this lane's results: entries 2 n0 and 2 n0 + 1 of its column
#pragma unroll
        for (int c = 0; c < CB; c++)
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const uint32_t o0 = base[c] ^ (uint32_t)fw->offp[p][2 * n0], o1 = base[c] ^ (uint32_t)fw->offp[p][2 * n0 + 1];
                t_psi[o0] = {ap[c][p][0], ap[c][p][1]};
                t_psi[o1] = {ap[c][p][2], ap[c][p][3]};
            }
        __syncthreads();
    }
}

// psi <- M psi and dpsi <- M dpsi + N psi over the windows [w0, w1) (the Hessian-vector product's
// forward pass); fragments [window][16][lane], M's 8 then N's 8
template <int NW>
__device__ inline void fwd_windows_hvp(const FWin* fwin, const float* fwaf, int w0, int w1, cpx<float>* t_psi,
                                       cpx<float>* t_dpsi, int lane, int wv)
{
    constexpr int CB = 16 / NW;
    const int n0 = lane >> 4, col = lane & 15;
    for (int w = w0; w < w1; w++) {
        const FWin* fw = fwin + w;
        float am[8], an[8];
#pragma unroll
        for (int f = 0; f < 8; f++) {
            am[f] = fwaf[((size_t)w * 16 + f) * 64 + lane];
            an[f] = fwaf[((size_t)w * 16 + 8 + f) * 64 + lane];
        }
        const uint32_t oj = (uint32_t)fw->offj[col];
        uint32_t base[CB];
        cpx<float> x[CB][2][2], xd[CB][2][2];
#pragma unroll
        for (int c = 0; c < CB; c++) {
            base[c] = oj ^ (uint32_t)fw->offcb[wv * CB + c];
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const uint32_t o = base[c] ^ (uint32_t)fw->offp[p][4 * h + n0];
                    x[c][p][h] = t_psi[o];
                    xd[c][p][h] = t_dpsi[o];
                }
        }
        v4f ap[CB][2], ad[CB][2];
#pragma unroll
        for (int c = 0; c < CB; c++)
#pragma unroll
            for (int p = 0; p < 2; p++) {
                ap[c][p] = v4f{0.f, 0.f, 0.f, 0.f};
                ad[c][p] = v4f{0.f, 0.f, 0.f, 0.f};
            }
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
#pragma unroll
            for (int c = 0; c < CB; c++)
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const cpx<float>& xv = x[c][p][kk >> 1];
                    const cpx<float>& dv = xd[c][p][kk >> 1];
                    const float bx = (kk & 1) ? xv.im : xv.re, bd = (kk & 1) ? dv.im : dv.re;
                    ap[c][p] = __builtin_amdgcn_mfma_f32_16x16x4f32(am[4 * p + kk], bx, ap[c][p], 0, 0, 0);
                    ad[c][p] = __builtin_amdgcn_mfma_f32_16x16x4f32(am[4 * p + kk], bd, ad[c][p], 0, 0, 0);
                    ad[c][p] = __builtin_amdgcn_mfma_f32_16x16x4f32(an[4 * p + kk], bx, ad[c][p], 0, 0, 0);
                }
#pragma unroll
        for (int c = 0; c < CB; c++)
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const uint32_t o0 = base[c] ^ (uint32_t)fw->offp[p][2 * n0], o1 = base[c] ^ (uint32_t)fw->offp[p][2 * n0 + 1];
                t_psi[o0] = {ap[c][p][0], ap[c][p][1]};
                t_psi[o1] = {ap[c][p][2], ap[c][p][3]};
                t_dpsi[o0] = {ad[c][p][0], ad[c][p][1]};
                t_dpsi[o1] = {ad[c][p][2], ad[c][p][3]};
            }
        __syncthreads();
    }
}

typedef double v4d __attribute__((ext_vector_type(4)));

// FP64: psi <- M psi (and with HVP dpsi <- M dpsi + N psi) over the windows [0, nwin) of a 12-qubit
// column in LDS laid out by mat_swz (struct DWin; fragments [window][8][lane], or [window][16][lane]
// with HVP, forward_windows64)
template <int NW, bool HVP>
__device__ inline void fwd_windows64(const DWin* wins, const double* af, int nwin, cpx<double>* psi, cpx<double>* dpsi,
                                     int lane, int wv)
{
    constexpr int CB = 16 / NW;
    constexpr int NF = HVP ? 16 : 8;
    const int n0 = lane >> 4, col = lane & 15;
    for (int w = 0; w < nwin; w++) {
        const DWin* dw = wins + w;
        double am[8], an[8];
#pragma unroll
        for (int f = 0; f < 8; f++) {
            am[f] = af[((size_t)w * NF + f) * 64 + lane];
            if constexpr (HVP) an[f] = af[((size_t)w * NF + 8 + f) * 64 + lane];
        }
        const uint32_t oj = (uint32_t)dw->offj[col];
        uint32_t base[CB];
        cpx<double> x[CB][2][2], xd[CB][2][2];
#pragma unroll
        for (int c = 0; c < CB; c++) {
            base[c] = oj ^ (uint32_t)dw->offcb[wv * CB + c];
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const uint32_t o = base[c] ^ (uint32_t)dw->offp[p][4 * h + n0];
                    x[c][p][h] = psi[o];
                    if constexpr (HVP) xd[c][p][h] = dpsi[o];
                }
        }
        v4d ap[CB][2], ad[CB][2];
#pragma unroll
        for (int c = 0; c < CB; c++)
#pragma unroll
            for (int p = 0; p < 2; p++) {
                ap[c][p] = v4d{0.0, 0.0, 0.0, 0.0};
                ad[c][p] = v4d{0.0, 0.0, 0.0, 0.0};
            }
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
#pragma unroll
            for (int c = 0; c < CB; c++)
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const cpx<double>& xv = x[c][p][kk >> 1];
                    const double bx = (kk & 1) ? xv.im : xv.re;
                    ap[c][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[4 * p + kk], bx, ap[c][p], 0, 0, 0);
                    if constexpr (HVP) {
                        const cpx<double>& dv = xd[c][p][kk >> 1];
                        ad[c][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[4 * p + kk], (kk & 1) ? dv.im : dv.re,
                                                                        ad[c][p], 0, 0, 0);
                        ad[c][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(an[4 * p + kk], bx, ad[c][p], 0, 0, 0);
                    }
                }
        // this lane's results: entries 2 n0 and 2 n0 + 1 of its column
#pragma unroll
        for (int c = 0; c < CB; c++)
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const uint32_t o0 = base[c] ^ (uint32_t)dw->offp[p][2 * n0], o1 = base[c] ^ (uint32_t)dw->offp[p][2 * n0 + 1];
                psi[o0] = {ap[c][p][0], ap[c][p][1]};
                psi[o1] = {ap[c][p][2], ap[c][p][3]};
                if constexpr (HVP) {
                    dpsi[o0] = {ad[c][p][0], ad[c][p][1]};
                    dpsi[o1] = {ad[c][p][2], ad[c][p][3]};
                }
            }
        __syncthreads();
    }
}

}  // namespace mg
