// elem3d.hpp -- stiffness of solid elements whose material varies in azimuth
// (the reference's "3D" path, SolidElement::displToStiff with mElem3D,
// SOLVER/src/core/element/SolidElement.cpp:404-443).  One workgroup owns one
// element and keeps everything in LDS:
//
//   for each row i of GLL points (5 passes of 5 points):
//     strain spectrum of the 5 points, all modes      Gradient::computeGrad6 + CrdTransTIsoSolid
//     C2R of 30 columns (packed as 15 complex FFTs)   FieldFFT::transformF2P, SolverFFTW_N6::computeC2R
//     stress at every azimuthal sample               Isotropic3D / TransverselyIsotropic3D / Anisotropic3D
//     R2C x 1/Nr                                      FieldFFT::transformP2F, SolverFFTW_N6::computeR2C
//     rotate back, quadrature operands, accumulate    CrdTransTIsoSolid::transformRTZ_SPZ, Gradient::computeQuad6
//   subtract the element's stiffness from its 25 points (mesh colouring: no races)
//
// The FFT is a mixed-radix (2,3,4,5,7,11,13) Stockham transform in LDS, done
// in place: every stage reads all its inputs into registers, synchronises,
// then writes.  Two real columns share one complex transform (z = x + i y).
#pragma once
#include "common.hpp"

namespace axg {

constexpr int E3D_THREADS = 512;
constexpr int E3D_STAGE_MODES = 64;  // modes of displacement staged at once
constexpr int E3D_MAX_NR = 272;      // 15 * Nr / radix <= threads * butterflies-per-thread
constexpr int E3D_MAX_FACTORS = 12;


struct Elem3DArgs {
    const int32_t *elems;
    int32_t nelem;
    int32_t maxM;
    const int32_t *elem_flags;
    const int32_t *elem_nr;
    const float *geom;
    const float *tiso;
    const float *mat3d;          // per element: [i][param][n][j]
    const int64_t *mat3d_off;    // [nelem]
    const int32_t *ep_row;
    const int32_t *ep_amax;
    const float2 *displ;
    float2 *stiff;
    const float2 *twiddle;       // W_N^t = exp(-2 pi i t / N), concatenated per N
    const int32_t *tw_off;       // [N] offset of the table of N
    const int32_t *factors;      // [N][8]: stage counts of radices 4, 2, 3, 5, 7, 11, 13
};

// ---------------------------------------------------------------- FFT stage

template <int R>
struct MaxB {
    static constexpr int v = (R == 2) ? 4 : (R == 3) ? 3 : (R <= 7) ? 2 : 1;  // 15 Nr / R <= B * threads for Nr <= 272
};

// one Stockham stage on nseq sequences of length N stored [seq][N] in `buf`:
// for butterfly j: inputs x[j + r N/R] * W_{Ns R}^{(j % Ns) r}, radix-R DFT,
// outputs y[(j / Ns) Ns R + (j % Ns) + r Ns].  sign = -1 forward, +1 inverse.
// LDS offsets of the W_R^m tables of the odd radices inside the twiddle area
template <int R>
constexpr int rootOff() {
    return (R == 3) ? 0 : (R == 5) ? 3 : (R == 7) ? 8 : (R == 11) ? 15 : 26;
}
constexpr int kRootTable = 39;

// One Stockham stage.  `tw` points into LDS: W_N^t for t < N, followed by the
// radix tables (loop-invariant constants would otherwise be hoisted into
// registers across the whole element loop).
template <int R>
__device__ __forceinline__ void fft_stage(float2 *buf, int nseq, int N, int Ns, int sign, const float2 *tw) {
    constexpr int B = MaxB<R>::v;
    const int nb = N / R;
    const int total = nseq * nb;
    const uint32_t mnb = udiv_magic(nb), mNs = udiv_magic(Ns);
    const int wmul = nb / Ns;  // W_{Ns R}^{jm} = W_N^{jm N / (Ns R)}
    float2 v[B][R];
    int dst[B];
#pragma unroll
    for (int q = 0; q < B; q++) {
        int t = opaque((int)threadIdx.x) + q * E3D_THREADS;
        dst[q] = -1;
        if (t < total) {
            int s = udiv(t, nb, mnb), j = t - s * nb;
            int jq = udiv(j, Ns, mNs), jm = j - jq * Ns;
            const float2 *x = buf + s * N;
            const int wstep = jm * wmul;
#pragma unroll
            for (int r = 0; r < R; r++) {
                float2 xv = x[j + r * nb];
                if (r == 0) {
                    v[q][r] = xv;
                } else {
                    float2 w = tw[wstep * r];  // the table holds the forward sign
                    if (sign > 0) w.y = -w.y;
                    v[q][r] = cmul(xv, w);
                }
            }
            dst[q] = s * N + jq * Ns * R + jm;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < B; q++) {
        if (dst[q] < 0) continue;
        float2 y[R];
        if constexpr (R == 2) {
            y[0] = cadd(v[q][0], v[q][1]);
            y[1] = csub(v[q][0], v[q][1]);
        } else if constexpr (R == 4) {
            float2 a0 = cadd(v[q][0], v[q][2]), a1 = csub(v[q][0], v[q][2]);
            float2 b0 = cadd(v[q][1], v[q][3]), b1 = csub(v[q][1], v[q][3]);
            // multiply b1 by -i (forward) or +i (inverse)
            float2 b1r = (sign < 0) ? make_float2(b1.y, -b1.x) : make_float2(-b1.y, b1.x);
            y[0] = cadd(a0, b0);
            y[2] = csub(a0, b0);
            y[1] = cadd(a1, b1r);
            y[3] = csub(a1, b1r);
        } else {
            // radix-R DFT y_k = sum_r v_r W_R^{rk}
#pragma unroll
            for (int k = 0; k < R; k++) {
                float2 acc = v[q][0];
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = tw[N + rootOff<R>() + (r * k) % R];
                    if (sign > 0) w.y = -w.y;
                    acc = cadd(acc, cmul(v[q][r], w));
                }
                y[k] = acc;
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) buf[dst[q] + r * Ns] = y[r];
    }
    __syncthreads();
}

// the full transform: fac = stage counts [n4, n2, n3, n5, n7, n11, n13] of N,
// applied in this order (one call site per radix keeps register use per stage)
__device__ __forceinline__ void fft_inplace(float2 *buf, int nseq, int N, int sign, const int32_t *fac,
                                            const float2 *tw) {
    int Ns = 1;
#pragma unroll 1
    for (int f = 0; f < fac[0]; f++) {
        fft_stage<4>(buf, nseq, N, Ns, sign, tw);
        Ns *= 4;
    }
#pragma unroll 1
    for (int f = 0; f < fac[1]; f++) {
        fft_stage<2>(buf, nseq, N, Ns, sign, tw);
        Ns *= 2;
    }
#pragma unroll 1
    for (int f = 0; f < fac[2]; f++) {
        fft_stage<3>(buf, nseq, N, Ns, sign, tw);
        Ns *= 3;
    }
#pragma unroll 1
    for (int f = 0; f < fac[3]; f++) {
        fft_stage<5>(buf, nseq, N, Ns, sign, tw);
        Ns *= 5;
    }
#pragma unroll 1
    for (int f = 0; f < fac[4]; f++) {
        fft_stage<7>(buf, nseq, N, Ns, sign, tw);
        Ns *= 7;
    }
    if (fac[5]) {
        fft_stage<11>(buf, nseq, N, Ns, sign, tw);
        Ns *= 11;
    }
    if (fac[6]) fft_stage<13>(buf, nseq, N, Ns, sign, tw);
}

// ---------------------------------------------------------------- stress
template <int MAT>
__device__ __forceinline__ void stress3d(const float *__restrict__ mp, int Nr, int n, int j, const float e[6],
                                         float s[6]) {
    // parameter p of sample n, point j (within the pass row): mp[(p * Nr + n) * 5 + j]
    auto P = [&](int p) { return mp[((int64_t)p * Nr + n) * 5 + j]; };
    if (MAT == 0) {
        // Isotropic3D::strainToStress (Isotropic3D.cpp:10-28)
        float lam = P(0), mu = P(1), mu2 = 2.f * mu;
        float sii = lam * (e[0] + e[1] + e[2]);
        s[0] = sii + mu2 * e[0];
        s[1] = sii + mu2 * e[1];
        s[2] = sii + mu2 * e[2];
        s[3] = mu * e[3];
        s[4] = mu * e[4];
        s[5] = mu * e[5];
    } else if (MAT == 1) {
        // TransverselyIsotropic3D::strainToStress (TransverselyIsotropic3D.cpp:10-28)
        float A = P(0), Cc = P(1), F = P(2), L = P(3), N = P(4), N2 = 2.f * N;
        float e01 = e[0] + e[1];
        float tmp = A * e01 + F * e[2];
        s[0] = tmp - N2 * e[1];
        s[1] = tmp - N2 * e[0];
        s[2] = Cc * e[2] + F * e01;
        s[3] = L * e[3];
        s[4] = L * e[4];
        s[5] = N * e[5];
    } else {
        // Anisotropic3D::strainToStress (Anisotropic3D.cpp:10-54)
        const int idx[6][6] = {{0, 1, 2, 3, 4, 5},     {1, 6, 7, 8, 9, 10},    {2, 7, 11, 12, 13, 14},
                               {3, 8, 12, 15, 16, 17}, {4, 9, 13, 16, 18, 19}, {5, 10, 14, 17, 19, 20}};
#pragma unroll
        for (int a = 0; a < 6; a++) {
            float acc = 0.f;
#pragma unroll
            for (int b = 0; b < 6; b++) acc += P(idx[a][b]) * e[b];
            s[a] = acc;
        }
    }
}

// same, with the parameters already in registers: P[p] of one (sample, point)
template <int MAT>
__device__ __forceinline__ void stress3d_reg(const float *P, const float e[6], float s[6]) {
    if (MAT == 0) {
        float lam = P[0], mu = P[1], mu2 = 2.f * mu;
        float sii = lam * (e[0] + e[1] + e[2]);
        s[0] = sii + mu2 * e[0];
        s[1] = sii + mu2 * e[1];
        s[2] = sii + mu2 * e[2];
        s[3] = mu * e[3];
        s[4] = mu * e[4];
        s[5] = mu * e[5];
    } else {
        float A = P[0], Cc = P[1], F = P[2], L = P[3], N = P[4], N2 = 2.f * N;
        float e01 = e[0] + e[1];
        float tmp = A * e01 + F * e[2];
        s[0] = tmp - N2 * e[1];
        s[1] = tmp - N2 * e[0];
        s[2] = Cc * e[2] + F * e01;
        s[3] = L * e[3];
        s[4] = L * e[4];
        s[5] = N * e[5];
    }
}

template <int MAT>
struct MatParams {
    static constexpr int v = (MAT == 0) ? 2 : (MAT == 1) ? 5 : 21;
};

// ---------------------------------------------------------------- kernel
// Work split: item (j, alpha) -> thread tid + q * T (q < QB).  The same thread
// owns that item in every pass, so the element's stiffness accumulates in its
// registers: acc[q][ip][c] = f_c(ip, j) of mode alpha.
template <int MAT, bool TISO>
__global__ __launch_bounds__(E3D_THREADS) void solid3d_kernel(Elem3DArgs a) {
    extern __shared__ float2 lds[];
    constexpr int T = E3D_THREADS;
    constexpr int QB = 2;  // 5 * M <= QB * T
    const int tid = threadIdx.x;
    const int Mmax = a.maxM;
    float2 *buf = lds;                  // spectral [30][M] / sequences [15][Nr]
    float2 *stg = buf + 30 * Mmax;      // [E3D_STAGE_MODES][3][25]
    float2 *twl = stg + 75 * min(E3D_STAGE_MODES, Mmax);  // W_Nr^t, t < Nr
    const int64_t e = a.elems[blockIdx.x];
    const int Nr = a.elem_nr[e];
    const int Nu = Nr / 2;
    const int M = Nu + 1;
    const int nyq = (Nr % 2 == 0) ? 1 : 0;
    const bool axial = (a.elem_flags[e] & 2) != 0;
    const float *Gx = cG[axial ? 1 : 0];
    const float *Ge = cG[0];
    const int32_t *fac = a.factors + Nr * 8;
    for (int t = tid; t < Nr; t += E3D_THREADS) twl[t] = a.twiddle[a.tw_off[Nr] + t];
    for (int t = tid; t < kRootTable; t += E3D_THREADS) twl[Nr + t] = a.twiddle[t];  // radix tables
    const float2 *tw = twl;  // read by every FFT stage; visible after the first barrier
    const float invNr = 1.f / (float)Nr;
    const uint32_t mM = udiv_magic(M);
    constexpr int NPAR = MatParams<MAT>::v;

    float2 acc[QB][5][3];
#pragma unroll
    for (int q = 0; q < QB; q++)
#pragma unroll
        for (int ip = 0; ip < 5; ip++)
#pragma unroll
            for (int c = 0; c < 3; c++) acc[q][ip][c] = czero();

#pragma unroll 1
    for (int i = 0; i < 5; i++) {
        const int tl = opaque(tid);
        // ---------------- strain spectrum of points (i, 0..4), modes 0..Nu
#pragma unroll 1
        for (int a0 = 0; a0 < M; a0 += E3D_STAGE_MODES) {
            const int nch = min(E3D_STAGE_MODES, M - a0);
            __syncthreads();
            for (int t = tl; t < nch * 25; t += T) {
                int q = t / 25, k = t - q * 25, al = a0 + q;
                int row = a.ep_row[e * 25 + k];
                bool ok = al <= a.ep_amax[e * 25 + k];
                const float2 *src = a.displ + ((int64_t)row + al) * 3;
#pragma unroll
                for (int c = 0; c < 3; c++) stg[(q * 3 + c) * 25 + k] = ok ? src[c] : czero();
            }
            __syncthreads();
            for (int t = tl; t < nch * 5; t += T) {
                int q = t / 5, j = t - q * 5, al = a0 + q;
                float2 e6[6];
                if (nyq && al == Nu) {
#pragma unroll
                    for (int c = 0; c < 6; c++) e6[c] = czero();
                } else {
                    const float2 *U = stg + q * 75;  // U[c*25 + m*5 + n]
                    PointGeom g = load_geom(a.geom, e, i * 5 + j);
                    float2 GU[3], UG[3];
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        float2 x = czero(), y = czero();
#pragma unroll
                        for (int m = 0; m < 5; m++) {
                            float gx = Gx[m * 5 + i], ge = Ge[m * 5 + j];
                            float2 um = U[c * 25 + m * 5 + j], un = U[c * 25 + i * 5 + m];
                            x.x += gx * um.x;
                            x.y += gx * um.y;
                            y.x += un.x * ge;
                            y.y += un.y * ge;
                        }
                        GU[c] = x;
                        UG[c] = y;
                    }
                    float fa = (float)al;
                    float2 u0 = U[0 * 25 + i * 5 + j], u1 = U[1 * 25 + i * 5 + j], u2 = U[2 * 25 + i * 5 + j];
                    float2 v0 = cadd(u0, cimul(fa, u1)), v1 = csub(cimul(fa, u0), u1), v2 = cimul(fa, u2);
                    e6[0] = cadd(cscale(g.dzdeta, GU[0]), cscale(g.dzdxii, UG[0]));
                    e6[1] = cscale(g.inv_s, v0);
                    e6[2] = cadd(cscale(g.dsdeta, GU[2]), cscale(g.dsdxii, UG[2]));
                    e6[3] = cadd(cadd(cscale(g.dsdeta, GU[1]), cscale(g.dsdxii, UG[1])), cscale(g.inv_s, v2));
                    e6[4] = cadd(cadd(cscale(g.dsdeta, GU[0]), cscale(g.dsdxii, UG[0])),
                                 cadd(cscale(g.dzdeta, GU[2]), cscale(g.dzdxii, UG[2])));
                    e6[5] = cadd(cadd(cscale(g.dzdeta, GU[1]), cscale(g.dzdxii, UG[1])), cscale(g.inv_s, v1));
                    if (axial && i == 0) {
                        e6[1] = cadd(e6[1], cscale(g.dzdeta, cadd(GU[0], cimul(fa, GU[1]))));
                        e6[5] = cadd(e6[5], cscale(g.dzdeta, csub(cimul(fa, GU[0]), GU[1])));
                        e6[3] = cadd(e6[3], cscale(g.dzdeta, cimul(fa, GU[2])));
                        if (al == 1) {
                            e6[1] = cadd(e6[1], cscale(g.dzdxii, cadd(UG[0], cimul(fa, UG[1]))));
                            e6[5] = cadd(e6[5], cscale(g.dzdxii, csub(cimul(fa, UG[0]), UG[1])));
                        }
                    }
                    if (TISO) {
                        const float *tp = a.tiso + e * 100 + i * 5 + j;
                        tiso_spz_rtz(e6, tp[0], tp[25], tp[50], tp[75]);
                    }
                }
#pragma unroll
                for (int c = 0; c < 6; c++) buf[(c * 5 + j) * M + al] = e6[c];
            }
        }
        __syncthreads();

        // material of this pass row, issued before the C2R transform so that the
        // loads overlap it (iso / TI: 2 or 5 parameters per sample and point)
        constexpr int SB = 3;  // 5 * Nr <= SB * T
        constexpr bool PREF = NPAR <= 5;
        float pm[PREF ? SB : 1][PREF ? NPAR : 1];
        const float *mp = a.mat3d + a.mat3d_off[e] + (int64_t)i * NPAR * Nr * 5;
        if constexpr (PREF) {
#pragma unroll
            for (int q = 0; q < SB; q++) {
                int t = tl + q * T;
                if (t < 5 * Nr) {
                    int n = t / 5, j = t - n * 5;
#pragma unroll
                    for (int p = 0; p < NPAR; p++) pm[q][p] = mp[((int64_t)p * Nr + n) * 5 + j];
                }
            }
        }

        // ---------------- pack 30 spectra into 15 Hermitian-extended sequences
        {
            constexpr int PB = 4;  // 15 * M <= PB * T
            float2 X[PB], Y[PB];
#pragma unroll
            for (int q = 0; q < PB; q++) {
                int t = tl + q * T;
                if (t < 15 * M) {
                    int b = udiv(t, M, mM), k = t - b * M;
                    X[q] = buf[(2 * b) * M + k];
                    Y[q] = buf[(2 * b + 1) * M + k];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < PB; q++) {
                int t = tl + q * T;
                if (t >= 15 * M) continue;
                int b = udiv(t, M, mM), k = t - b * M;
                float2 *z = buf + b * Nr;
                if (k == 0 || 2 * k == Nr) {
                    // DC and Nyquist: the imaginary parts are ignored (FFTW c2r)
                    z[k] = make_float2(X[q].x, Y[q].x);
                } else {
                    // Z_k = X_k + i Y_k, Z_{N-k} = conj(X_k) + i conj(Y_k)
                    z[k] = make_float2(X[q].x - Y[q].y, X[q].y + Y[q].x);
                    z[Nr - k] = make_float2(X[q].x + Y[q].y, -X[q].y + Y[q].x);
                }
            }
            __syncthreads();
        }
        fft_inplace(buf, 15, Nr, +1, fac, tw);

        // ---------------- stress at every (sample n, point j)
        {
#pragma unroll
            for (int q = 0; q < SB; q++) {
                int t = tl + q * T;
                if (t >= 5 * Nr) continue;
                int n = t / 5, j = t - n * 5;
                float ev[6], sv[6];
#pragma unroll
                for (int c = 0; c < 6; c++) {
                    int col = c * 5 + j;
                    float2 z = buf[(col >> 1) * Nr + n];
                    ev[c] = (col & 1) ? z.y : z.x;
                }
                if constexpr (PREF) stress3d_reg<MAT>(pm[q], ev, sv);
                else stress3d<MAT>(mp, Nr, n, j, ev, sv);
#pragma unroll
                for (int c = 0; c < 6; c++) {
                    int col = c * 5 + j;
                    float2 *z = buf + (col >> 1) * Nr + n;
                    if (col & 1) z->y = sv[c];
                    else z->x = sv[c];
                }
            }
            __syncthreads();
        }
        fft_inplace(buf, 15, Nr, -1, fac, tw);

        // ---------------- unpack: X_k = (Z_k + conj Z_{N-k}) / 2, Y_k = (Z_k - conj Z_{N-k}) / 2i, x 1/Nr
        {
            constexpr int PB = 4;
            float2 Z0[PB], Z1[PB];
#pragma unroll
            for (int q = 0; q < PB; q++) {
                int t = tl + q * T;
                if (t < 15 * M) {
                    int b = udiv(t, M, mM), k = t - b * M;
                    const float2 *z = buf + b * Nr;
                    Z0[q] = z[k];
                    Z1[q] = z[k == 0 ? 0 : Nr - k];
                }
            }
            __syncthreads();
            const float h = 0.5f * invNr;
#pragma unroll
            for (int q = 0; q < PB; q++) {
                int t = tl + q * T;
                if (t >= 15 * M) continue;
                int b = udiv(t, M, mM), k = t - b * M;
                float2 za = Z0[q], zb = Z1[q];
                buf[(2 * b) * M + k] = make_float2(h * (za.x + zb.x), h * (za.y - zb.y));
                buf[(2 * b + 1) * M + k] = make_float2(h * (za.y + zb.y), h * (zb.x - za.x));
            }
            __syncthreads();
        }

        // ---------------- rotate back, quadrature operands, accumulate in registers.
        // A thread reads the stress of its own (j, alpha) items and overwrites the
        // same LDS slots with its Y operands, so no barrier is needed in between.
        {
#pragma unroll
            for (int q = 0; q < QB; q++) {
                int t = tl + q * T;
                int al = t / 5, j = t - al * 5;
                if (t < 5 * M && !(nyq && al == Nu)) {
                    float2 s6[6];
#pragma unroll
                    for (int c = 0; c < 6; c++) s6[c] = buf[(c * 5 + j) * M + al];
                    if (TISO) {
                        const float *tp = a.tiso + e * 100 + i * 5 + j;
                        tiso_rtz_spz(s6, tp[0], tp[25], tp[50], tp[75]);
                    }
                    PointGeom g = load_geom(a.geom, e, i * 5 + j);
                    float nb = -(float)al;
                    float2 X[3], Y[3], gg[3];
                    X[0] = cadd(cscale(g.dzdeta, s6[0]), cscale(g.dsdeta, s6[4]));
                    X[1] = cadd(cscale(g.dzdeta, s6[5]), cscale(g.dsdeta, s6[3]));
                    X[2] = cadd(cscale(g.dzdeta, s6[4]), cscale(g.dsdeta, s6[2]));
                    Y[0] = cadd(cscale(g.dzdxii, s6[0]), cscale(g.dsdxii, s6[4]));
                    Y[1] = cadd(cscale(g.dzdxii, s6[5]), cscale(g.dsdxii, s6[3]));
                    Y[2] = cadd(cscale(g.dzdxii, s6[4]), cscale(g.dsdxii, s6[2]));
                    gg[0] = cadd(s6[1], cimul(nb, s6[5]));
                    gg[1] = csub(cimul(nb, s6[1]), s6[5]);
                    gg[2] = cimul(nb, s6[3]);
                    // own point: inv_s g
#pragma unroll
                    for (int ip = 0; ip < 5; ip++)
                        if (ip == i)
#pragma unroll
                            for (int c = 0; c < 3; c++) acc[q][ip][c] = cadd(acc[q][ip][c], cscale(g.inv_s, gg[c]));
                    // axial row 0: G_xi(:,0) dzdeta g has the shape of the X term,
                    // (dzdxii g) G^T_eta that of the Y term
                    if (axial && i == 0) {
#pragma unroll
                        for (int c = 0; c < 3; c++) X[c] = cadd(X[c], cscale(g.dzdeta, gg[c]));
                        if (al == 1) {
                            Y[0] = cadd(Y[0], cscale(g.dzdxii, gg[0]));
                            Y[1] = cadd(Y[1], cscale(g.dzdxii, gg[1]));
                        }
                    }
                    // X term: column j of every row.  For axial elements the W term
                    // uses G_xi(ip, 0) and the X term G_xi(ip, i); on row 0 both are
                    // G_xi(ip, 0), so they were merged above.
#pragma unroll
                    for (int c = 0; c < 3; c++)
#pragma unroll
                        for (int ip = 0; ip < 5; ip++)
                            acc[q][ip][c] = cadd(acc[q][ip][c], cscale(Gx[ip * 5 + i], X[c]));
#pragma unroll
                    for (int c = 0; c < 3; c++) buf[(c * 5 + j) * M + al] = Y[c];
                }
            }
            __syncthreads();
            // Y term: row i, f(i, j) += sum_m Y(i, m) G_eta(j, m)
#pragma unroll
            for (int q = 0; q < QB; q++) {
                int t = tl + q * T;
                int al = t / 5, j = t - al * 5;
                if (t < 5 * M && !(nyq && al == Nu)) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        float2 f = czero();
#pragma unroll
                        for (int m = 0; m < 5; m++) {
                            float ge = Ge[j * 5 + m];
                            float2 y = buf[(c * 5 + m) * M + al];
                            f.x += y.x * ge;
                            f.y += y.y * ge;
                        }
#pragma unroll
                        for (int ip = 0; ip < 5; ip++)
                            if (ip == i) acc[q][ip][c] = cadd(acc[q][ip][c], f);
                    }
                }
            }
        }
    }

    // ---------------- gatherStiffFromElement: stiff -= f on valid rows
#pragma unroll
    for (int q = 0; q < QB; q++) {
        int t = tid + q * T;
        int al = t / 5, j = t - al * 5;
        if (t >= 5 * M || (nyq && al == Nu)) continue;
#pragma unroll
        for (int ip = 0; ip < 5; ip++) {
            int k = ip * 5 + j;
            if (al > a.ep_amax[e * 25 + k]) continue;
            float2 *sp = a.stiff + ((int64_t)a.ep_row[e * 25 + k] + al) * 3;
#pragma unroll
            for (int c = 0; c < 3; c++) sp[c] = csub(sp[c], acc[q][ip][c]);
        }
    }
}

}  // namespace axg
