// pointfft.hpp -- point operators whose coefficients vary in azimuth, applied
// between Fourier and physical space at a point's own Nr:
//   Mass3D::computeAccel                SOLVER/src/core/point/mass/Mass3D.cpp:13-57
//   SFCoupling3D::coupleSolidToFluid /  SOLVER/src/core/point/solid_fluid/SFCoupling3D.cpp:9-55
//                 coupleFluidToSolid
// One wave per point (4 per 256-thread block, points strided over the grid),
// each wave with a private LDS area: twiddles of the point's Nr, and up to two
// interleaved sequences that carry the point's columns (two real columns per
// complex transform); the transforms are the elements' wave FFT.
#pragma once
#include "elem3d.hpp"
#include "points.hpp"

namespace axg {

constexpr int PF_WAVES = 4;
// float2 slots of one wave's LDS area for points of Nr <= maxNr: twiddles
// 2 Nr, sequences 2 Nr
// (+ the FFT stage table of the point's Nr, 48 slots)
constexpr int pfWaveLds(int maxNr) { return 4 * maxNr + E3D_MAX_STAGES * FS_N / 2; }

struct PointFFTArgs {
    int32_t n;                   // points in the list
    int32_t wave_lds;            // float2 slots of each wave's LDS area, pfWaveLds(largest Nr)
    const int32_t *items;        // Mass3D: compact points; coupling: SF3D entries
    const int32_t *pt_row;       // phase rows (Mass3D) / solid rows (coupling)
    const int32_t *pt_nu;
    const int32_t *pt_flags;
    const int32_t *pt_nr;        // Nr of each item's point
    const int64_t *coef_off;     // per item: first slice of its coefficients
    const float *coef;           // Mass3D: 1/mass [Nr]; coupling: [Nr][6] normals
    const int32_t *frow;         // coupling: fluid rows of the item
    float2 *field;               // Mass3D: stiff of the phase; coupling: solid stiff
    const float2 *sdispl;        // coupling: solid displacement
    float2 *fstiff;              // coupling: fluid stiff
    const float2 *twiddle;
    const int32_t *tw_off;
    const int32_t *plan;
    const int32_t *nstage;
};

// twiddles, and the stage table into pl (LDS) so that the stages
// read it with uniform LDS loads
__device__ __forceinline__ void pf_load_twiddles(const PointFFTArgs &a, int Nr, float2 *twf, float2 *twi,
                                                 int32_t *pl, int lane) {
    for (int t = lane; t < Nr; t += 64) {
        float2 v = a.twiddle[a.tw_off[Nr] + t];
        twf[t] = v;
        twi[t] = make_float2(v.x, -v.y);
    }
    const int32_t *gp = a.plan + Nr * E3D_MAX_STAGES * FS_N;
    for (int t = lane; t < E3D_MAX_STAGES * FS_N; t += 64) pl[t] = gp[t];
    wave_lds_fence();
}

// real spectrum X (row alpha of a column) into the Hermitian extension of
// sequence slot `b` (even b: real part, odd b: imaginary part) of [Nr][2]
__device__ __forceinline__ void pf_pack(float2 *seq, int Nr, int al, float2 X, int b) {
    float *z = reinterpret_cast<float *>(seq);
    const int s = b >> 1, im = b & 1;
    if (al == 0 || 2 * al == Nr) {
        z[(2 * al + s) * 2 + im] = X.x;  // imaginary parts of DC and Nyquist dropped
        return;
    }
    // column b contributes X to Z_al and conj(X) to Z_{Nr-al}, times i if odd
    if (im == 0) {
        z[(2 * al + s) * 2 + 0] += X.x;
        z[(2 * al + s) * 2 + 1] += X.y;
        z[(2 * (Nr - al) + s) * 2 + 0] += X.x;
        z[(2 * (Nr - al) + s) * 2 + 1] -= X.y;
    } else {
        z[(2 * al + s) * 2 + 0] -= X.y;
        z[(2 * al + s) * 2 + 1] += X.x;
        z[(2 * (Nr - al) + s) * 2 + 0] += X.y;
        z[(2 * (Nr - al) + s) * 2 + 1] += X.x;
    }
}

// spectrum of column b (row al) from the transformed sequences, times 1/Nr
__device__ __forceinline__ float2 pf_unpack(const float2 *seq, int Nr, int al, int b) {
    const int s = b >> 1;
    const float h = 0.5f / (float)Nr;
    float2 za = seq[2 * al + s], zb = seq[2 * (al == 0 ? 0 : Nr - al) + s];
    return (b & 1) ? make_float2(h * (za.y + zb.y), h * (zb.x - za.x))
                   : make_float2(h * (za.x + zb.x), h * (za.y - zb.y));
}

__device__ __forceinline__ void pf_clear(float2 *seq, int Nr, int lane) {
    for (int t = lane; t < 2 * Nr; t += 64) seq[t] = czero();
    wave_lds_fence();
}

// Mass3D: stiff -> mask -> C2R -> x 1/m(phi) -> R2C -> stiff (the Newmark
// kernel then masks again and scales by 1, SolidPoint.cpp:23-38).
// OCEAN (MassOcean3D::computeAccel, MassOcean3D.cpp:19-53): the coefficients
// are rows {1/m, n_s, n_phi, n_z} of each slice, n the unit surface normal times
// sqrt(m_o / (m (m + m_o))), and the product is a = s / m - (s . n) n, which
// divides the normal part by m + m_o and the tangential part by m.
template <int NCOMP, bool OCEAN>
__global__ __launch_bounds__(256) void mass3d_kernel(PointFFTArgs a) {
    extern __shared__ float2 lds[];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    float2 *twf = lds + w * a.wave_lds;
    for (int it = blockIdx.x * PF_WAVES + w; it < a.n; it += gridDim.x * PF_WAVES) {
        const int q = a.items[it];
        const int Nr = a.pt_nr[it], nu = a.pt_nu[q], fl = a.pt_flags[q];
        float2 *twi = twf + Nr, *seq = twi + Nr;
        int32_t *plan = reinterpret_cast<int32_t *>(seq + 2 * Nr);
        pf_load_twiddles(a, Nr, twf, twi, plan, lane);
        pf_clear(seq, Nr, lane);
        float2 *st = a.field + (int64_t)a.pt_row[q] * NCOMP;
        for (int al = lane; al <= nu; al += 64) {
            float2 v[3] = {czero(), czero(), czero()};
            for (int c = 0; c < NCOMP; c++) v[c] = st[al * NCOMP + c];
            if (NCOMP == 3) {
                mask_solid_row(v, al, nu, fl);
            } else {
                if (al == 0) v[0].y = 0.f;
                if (((fl & 1) && al >= 1) || ((fl & 2) && al == nu)) v[0] = czero();
            }
            for (int c = 0; c < NCOMP; c++) pf_pack(seq, Nr, al, v[c], c);
        }
        wave_lds_fence();
        wfft<E3D_MAX_NR, 2>(seq, Nr, +1, plan, a.nstage[Nr], twi, lane);
        const float *im = a.coef + a.coef_off[it];
        for (int n = lane; n < Nr; n += 64) {
            float2 z0 = seq[2 * n], z1 = seq[2 * n + 1];
            if (OCEAN) {
                const float4 q = *reinterpret_cast<const float4 *>(im + 4 * n);
                const float fn = z0.x * q.y + z0.y * q.z + z1.x * q.w;
                seq[2 * n] = make_float2(z0.x * q.x - fn * q.y, z0.y * q.x - fn * q.z);
                seq[2 * n + 1] = make_float2(z1.x * q.x - fn * q.w, 0.f);
            } else {
                const float m = im[n];
                seq[2 * n] = make_float2(m * z0.x, m * z0.y);
                seq[2 * n + 1] = make_float2(m * z1.x, m * z1.y);
            }
        }
        wave_lds_fence();
        wfft<E3D_MAX_NR, 2>(seq, Nr, -1, plan, a.nstage[Nr], twf, lane);
        for (int al = lane; al <= nu; al += 64)
            for (int c = 0; c < NCOMP; c++) st[al * NCOMP + c] = pf_unpack(seq, Nr, al, c);
        wave_lds_fence();
    }
}

// SFCoupling3D, in the order of SolidFluidPoint::coupleSolidFluid
// (SolidFluidPoint.cpp:106-110): fluid stiff += R2C(n . C2R(u_solid)), then
// solid stiff -= R2C((n / m_f) C2R(fluid stiff)).
__global__ __launch_bounds__(256) void couple3d_kernel(PointFFTArgs a) {
    extern __shared__ float2 lds[];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    float2 *twf = lds + w * a.wave_lds;
    for (int it = blockIdx.x * PF_WAVES + w; it < a.n; it += gridDim.x * PF_WAVES) {
        const int Nr = a.pt_nr[it], nu = Nr / 2;
        float2 *twi = twf + Nr, *seq = twi + Nr;
        int32_t *plan = reinterpret_cast<int32_t *>(seq + 2 * Nr);
        const float *nrm = a.coef + a.coef_off[it] * 6;
        const float2 *ud = a.sdispl + (int64_t)a.pt_row[it] * 3;
        float2 *ss = a.field + (int64_t)a.pt_row[it] * 3;
        float2 *fs = a.fstiff + a.frow[it];
        pf_load_twiddles(a, Nr, twf, twi, plan, lane);
        // solid -> fluid
        pf_clear(seq, Nr, lane);
        for (int al = lane; al <= nu; al += 64)
            for (int c = 0; c < 3; c++) pf_pack(seq, Nr, al, ud[al * 3 + c], c);
        wave_lds_fence();
        wfft<E3D_MAX_NR, 2>(seq, Nr, +1, plan, a.nstage[Nr], twi, lane);
        for (int n = lane; n < Nr; n += 64) {
            float2 z0 = seq[2 * n], z1 = seq[2 * n + 1];
            const float *nn = nrm + n * 6;
            seq[2 * n] = make_float2(nn[0] * z0.x + nn[1] * z0.y + nn[2] * z1.x, 0.f);
            seq[2 * n + 1] = czero();
        }
        wave_lds_fence();
        wfft<E3D_MAX_NR, 2>(seq, Nr, -1, plan, a.nstage[Nr], twf, lane);
        for (int al = lane; al <= nu; al += 64) fs[al] = cadd(fs[al], pf_unpack(seq, Nr, al, 0));
        wave_lds_fence();
        // fluid -> solid (with the updated fluid stiffness)
        pf_clear(seq, Nr, lane);
        for (int al = lane; al <= nu; al += 64) pf_pack(seq, Nr, al, fs[al], 0);
        wave_lds_fence();
        wfft<E3D_MAX_NR, 2>(seq, Nr, +1, plan, a.nstage[Nr], twi, lane);
        for (int n = lane; n < Nr; n += 64) {
            const float y = seq[2 * n].x;
            const float *nn = nrm + n * 6;
            seq[2 * n] = make_float2(nn[3] * y, nn[4] * y);
            seq[2 * n + 1] = make_float2(nn[5] * y, 0.f);
        }
        wave_lds_fence();
        wfft<E3D_MAX_NR, 2>(seq, Nr, -1, plan, a.nstage[Nr], twf, lane);
        for (int al = lane; al <= nu; al += 64)
            for (int c = 0; c < 3; c++) ss[al * 3 + c] = csub(ss[al * 3 + c], pf_unpack(seq, Nr, al, c));
        wave_lds_fence();
    }
}

}  // namespace axg
