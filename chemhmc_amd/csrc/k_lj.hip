// Lennard-Jones sweep over the 27 neighbour cells: the replacement of Tag_force_binning
// (code/modules/identical_particles.cpp:567-639), Tag_potential_binning (:395-454) and the kick of
// update_momenta (:676-694), with a pair virial the reference does not have.
//
// One 64-lane wavefront per i-cell (workgroup = one wave, no cross-wave barriers):
//   1. lanes 0..26 resolve the 27 neighbour cells (with the reference's explicit image shift when
//      the neighbour index wraps, :584-600) and a wave prefix sum gives the concatenated j-list;
//   2. the j-list is staged into LDS, row by row (3 consecutive cells per (dy,dz)), with the shift
//      folded into the staged coordinates, so the inner loop reads one broadcast LDS entry per pair;
//   3. lanes are IS i-slots x (64/IS) j-groups: lane = g*IS + ii takes i-atom ii of the current
//      i-chunk and every (64/IS)-th j entry; a butterfly over the j-groups sums the partials.
// FP32 path: coordinates relative to the cell origin (|d| < 2 cell sides), per-atom FP32 sums,
//   +grad U stored and/or the FP64 momenta kicked in place (p -= (dt*beta)*f).
// FP64 path: coordinates relative to the i-cell origin in FP64; optional per-atom energy/virial
//   sums reduced deterministically per cell and per chain (last-arriving workgroup sums in order).
// Self-exclusion is r^2 > 0: the staged copy of atom i is bitwise equal to its own coordinates.
#include "chmc_internal.hpp"

#pragma clang fp contract(off)   // the kick must round like the oracle's p -= dt*c*f

namespace chmc {

__device__ __forceinline__ int xcd_remap(int b, int nblocks)
{
    // blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of cells
    // so that neighbouring cells share one L2.
    const int per = nblocks >> 3;
    return (b & 7) * per + (b >> 3);
}

template <typename Real> struct Stage;
template <> struct Stage<float> { float4 q[kMaxJ]; };
template <> struct Stage<double> { double x[kMaxJ], y[kMaxJ], z[kMaxJ]; };

__device__ __forceinline__ float  rcp_(float r2) { return __builtin_amdgcn_rcpf(r2); }
__device__ __forceinline__ double rcp_(double r2) { return 1.0 / r2; }
__device__ __forceinline__ float  fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename Real, int IS, int MODE>
__global__ __launch_bounds__(64) void k_lj(DevView v, double kick_dt, int nblocks)
{
    constexpr int JG = 64 / IS;
    __shared__ Stage<Real> sj;
    __shared__ int seg_beg[27];
    __shared__ int seg_pref[28];
    __shared__ Real seg_sh[27][3];
    __shared__ int last_block;

    const int lane = threadIdx.x;
    const int c = xcd_remap(blockIdx.x, nblocks);
    if (c >= v.Ctot) return;
    const int ch = v.cell_chain[c];
    const ChainDev& C = v.chains[ch];
    const int nx = C.nbin[0], ny = C.nbin[1], nz = C.nbin[2];
    const int lc = c - C.cell_off;
    const int bz = lc / (nx * ny);
    const int by = (lc - bz * nx * ny) / nx;
    const int bx = lc - nx * (by + bz * ny);
    const int ibeg = v.offs[c];
    const int ni = v.offs[c + 1] - ibeg;
    const int N = v.Ntot;

    double cell_u = 0.0, cell_w = 0.0, cell_n = 0.0;
    if (ni > 0) {
        // ---- 1. neighbour segments -----------------------------------------------------------
        int cnt = 0;
        if (lane < 27) {
            const int dd[3] = { lane % 3 - 1, (lane / 3) % 3 - 1, lane / 9 - 1 };
            const int b3[3] = { bx, by, bz };
            const int n3[3] = { nx, ny, nz };
            int j3[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                j3[d] = b3[d] + dd[d];
                const bool wrapped = j3[d] < 0 || j3[d] >= n3[d];
                if (wrapped) j3[d] = (b3[d] + n3[d] + dd[d]) % n3[d];
                if constexpr (sizeof(Real) == 4) {
                    // staged = (x_j - o_j) + dd*s: the j cell's origin seen from the i cell's origin
                    seg_sh[lane][d] = (float)dd[d] * C.sizebin_f[d];
                } else {
                    // staged = x_j + wrap - o_i, wrap = dd*L only when the index wrapped
                    seg_sh[lane][d] = (wrapped ? dd[d] * C.L[d] : 0.0) - b3[d] * C.sizebin[d];
                }
            }
            const int jc = C.cell_off + j3[0] + nx * (j3[1] + j3[2] * ny);
            seg_beg[lane] = v.offs[jc];
            cnt = v.offs[jc + 1] - v.offs[jc];
        }
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane < 27) seg_pref[lane + 1] = incl;
        if (lane == 0) seg_pref[0] = 0;
        __syncthreads();
        const int nj = seg_pref[27];

        // ---- 2. staging of j-list entries [jlo, jhi) ------------------------------------------
        auto stage = [&](int jlo, int jhi) {
            for (int r = 0; r < 9; r++) {
                const int p0 = seg_pref[3 * r], p1 = seg_pref[3 * r + 1], p2 = seg_pref[3 * r + 2];
                const int p3 = seg_pref[3 * r + 3];
                const int lo = max(p0, jlo), hi = min(p3, jhi);
                for (int t = lo + lane; t < hi; t += 64) {
                    const int k = 3 * r + (t >= p1 ? 1 : 0) + (t >= p2 ? 1 : 0);
                    const int src = seg_beg[k] + (t - seg_pref[k]);
                    if constexpr (sizeof(Real) == 4) {
                        const float4 q = v.xs[src];
                        sj.q[t - jlo] = make_float4(q.x + seg_sh[k][0], q.y + seg_sh[k][1], q.z + seg_sh[k][2], 0.f);
                    } else {
                        const int a = v.perm[src];
                        sj.x[t - jlo] = v.x[a] + seg_sh[k][0];
                        sj.y[t - jlo] = v.x[N + a] + seg_sh[k][1];
                        sj.z[t - jlo] = v.x[2 * N + a] + seg_sh[k][2];
                    }
                }
            }
        };
        const bool single = nj <= kMaxJ;
        if (single) {
            stage(0, nj);
            __syncthreads();
        }

        // ---- 3. i-chunks x j-groups --------------------------------------------------------------
        const int ii = lane % IS, g = lane / IS;
        const Real rc2 = (sizeof(Real) == 4) ? (Real)C.rc2_f : (Real)C.rc2;
        const Real sig2 = (sizeof(Real) == 4) ? (Real)C.sigma2_f : (Real)C.sigma2;
        for (int i0 = 0; i0 < ni; i0 += IS) {
            const int i = i0 + ii;
            const bool ival = i < ni;
            const int ic = ival ? i : ni - 1;
            Real xi, yi, zi;
            int atom;
            if constexpr (sizeof(Real) == 4) {
                const float4 q = v.xs[ibeg + ic];
                xi = q.x; yi = q.y; zi = q.z;
                atom = __float_as_int(q.w);
            } else {
                atom = v.perm[ibeg + ic];
                xi = v.x[atom] + seg_sh[13][0];
                yi = v.x[N + atom] + seg_sh[13][1];
                zi = v.x[2 * N + atom] + seg_sh[13][2];
            }
            Real ax = 0, ay = 0, az = 0, au = 0, aw = 0;
            int an = 0;
            for (int jlo = 0; jlo < nj; jlo += kMaxJ) {
                const int jhi = min(nj, jlo + kMaxJ);
                if (!single) {
                    __syncthreads();
                    stage(jlo, jhi);
                    __syncthreads();
                }
                const int jn = jhi - jlo;
#pragma unroll 4
                for (int j = g; j < jn; j += JG) {
                    Real dx, dy, dz;
                    if constexpr (sizeof(Real) == 4) {
                        const float4 q = sj.q[j];
                        dx = xi - q.x; dy = yi - q.y; dz = zi - q.z;
                    } else {
                        dx = xi - sj.x[j]; dy = yi - sj.y[j]; dz = zi - sj.z[j];
                    }
                    const Real r2 = fma_(dz, dz, fma_(dy, dy, dx * dx));
                    const bool in = (r2 < rc2) && (r2 > (Real)0);
                    const Real inv = rcp_(r2);
                    const Real s2 = sig2 * inv;
                    const Real s6 = s2 * s2 * s2;
                    if constexpr ((MODE & LJ_FORCE) != 0) {
                        Real gq = s6 * ((Real)0.5 - s6) * inv;   // sr6 (0.5 - sr6) / r^2, :611-613
                        gq = in ? gq : (Real)0;
                        ax = fma_(gq, dx, ax);
                        ay = fma_(gq, dy, ay);
                        az = fma_(gq, dz, az);
                    }
                    if constexpr ((MODE & LJ_ENERGY) != 0) {
                        au += in ? s6 * (s6 - (Real)1) : (Real)0;
                        aw += in ? s6 * (s6 - (Real)0.5) : (Real)0;
                        an += in ? 1 : 0;
                    }
                }
            }
            // butterfly over the j-groups (lanes ii, ii+IS, ...): every lane ends with the i-atom total
#pragma unroll
            for (int m = IS; m < 64; m <<= 1) {
                if constexpr ((MODE & LJ_FORCE) != 0) {
                    ax += __shfl_xor(ax, m, 64);
                    ay += __shfl_xor(ay, m, 64);
                    az += __shfl_xor(az, m, 64);
                }
                if constexpr ((MODE & LJ_ENERGY) != 0) {
                    au += __shfl_xor(au, m, 64);
                    aw += __shfl_xor(aw, m, 64);
                    an += __shfl_xor(an, m, 64);
                }
            }
            if constexpr ((MODE & LJ_FORCE) != 0) {
                if (g == 0 && ival) {
                    Real f3[3];
                    if constexpr (sizeof(Real) == 4) {
                        f3[0] = C.eps48_f * ax; f3[1] = C.eps48_f * ay; f3[2] = C.eps48_f * az;
                    } else {
                        const double e48 = 48 * C.eps;
                        f3[0] = ax * e48; f3[1] = ay * e48; f3[2] = az * e48;
                    }
                    const double kdt = kick_dt * C.beta;     // dt * coeff_p, coeff_p = beta (:642-644)
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        if constexpr ((MODE & LJ_STOREF) != 0) {
                            if constexpr (sizeof(Real) == 4) v.f32[d * N + atom] = f3[d];
                            else v.f64[d * N + atom] = f3[d];
                        }
                        if constexpr ((MODE & LJ_KICK) != 0) {
                            const double fd = (double)f3[d];
                            v.p[d * N + atom] = v.p[d * N + atom] - kdt * fd;
                        }
                    }
                }
            }
            if constexpr ((MODE & LJ_ENERGY) != 0) {
                double eu = (g == 0 && ival) ? (double)au : 0.0;
                double ew = (g == 0 && ival) ? (double)aw : 0.0;
                int en = (g == 0 && ival) ? an : 0;
#pragma unroll
                for (int m = 1; m < IS; m <<= 1) {
                    eu += __shfl_xor(eu, m, 64);
                    ew += __shfl_xor(ew, m, 64);
                    en += __shfl_xor(en, m, 64);
                }
                cell_u += eu;
                cell_w += ew;
                cell_n += (double)en;
            }
        }
    }

    if constexpr ((MODE & LJ_ENERGY) != 0) {
        // per-cell partials, then the last workgroup of the chain sums them in cell order
        if (lane == 0) {
            v.cell_part[3 * c] = cell_u;
            v.cell_part[3 * c + 1] = cell_w;
            v.cell_part[3 * c + 2] = cell_n;
        }
        __threadfence();
        if (lane == 0) last_block = (atomicAdd(&v.ticket[ch], 1) == C.ncells - 1) ? 1 : 0;
        __syncthreads();
        if (last_block) {
            __threadfence();
            double su = 0.0, sw = 0.0, sn = 0.0;
            for (int q = lane; q < C.ncells; q += 64) {
                const double* cp = &v.cell_part[3 * (C.cell_off + q)];
                su += __hip_atomic_load(cp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                sw += __hip_atomic_load(cp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                sn += __hip_atomic_load(cp + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
#pragma unroll
            for (int m = 1; m < 64; m <<= 1) {
                su += __shfl_xor(su, m, 64);
                sw += __shfl_xor(sw, m, 64);
                sn += __shfl_xor(sn, m, 64);
            }
            if (lane == 0) {
                v.red[ch * RED_NSLOT + RED_NINT] = sn;                 // ordered interacting pairs
                v.red[ch * RED_NSLOT + RED_V] = 2 * C.eps * su;      // :392, 2 eps because pairs count twice
                v.red[ch * RED_NSLOT + RED_W] = 24 * C.eps * sw;     // sum_{i<j} 48 eps (sr12 - sr6/2)
                v.ticket[ch] = 0;
            }
        }
    }
}

template <typename Real, int MODE>
static hipError_t launch_is(const DevView& v, int islots, double kick_dt, hipStream_t s)
{
    const int nblocks = ((v.Ctot + 7) / 8) * 8;
    switch (islots) {
    case 4: hipLaunchKernelGGL((k_lj<Real, 4, MODE>), dim3(nblocks), dim3(64), 0, s, v, kick_dt, nblocks); break;
    case 16: hipLaunchKernelGGL((k_lj<Real, 16, MODE>), dim3(nblocks), dim3(64), 0, s, v, kick_dt, nblocks); break;
    default: hipLaunchKernelGGL((k_lj<Real, 8, MODE>), dim3(nblocks), dim3(64), 0, s, v, kick_dt, nblocks); break;
    }
    return hipGetLastError();
}

hipError_t launch_lj(const DevView& v, int mode, bool fp64, int islots, double kick_dt, hipStream_t s)
{
    if (!fp64) {
        switch (mode) {
        case LJ_FORCE | LJ_KICK: return launch_is<float, LJ_FORCE | LJ_KICK>(v, islots, kick_dt, s);
        case LJ_FORCE | LJ_STOREF: return launch_is<float, LJ_FORCE | LJ_STOREF>(v, islots, kick_dt, s);
        case LJ_FORCE | LJ_KICK | LJ_STOREF: return launch_is<float, LJ_FORCE | LJ_KICK | LJ_STOREF>(v, islots, kick_dt, s);
        default: return hipErrorInvalidValue;
        }
    }
    switch (mode) {
    case LJ_FORCE | LJ_KICK: return launch_is<double, LJ_FORCE | LJ_KICK>(v, islots, kick_dt, s);
    case LJ_FORCE | LJ_STOREF: return launch_is<double, LJ_FORCE | LJ_STOREF>(v, islots, kick_dt, s);
    case LJ_FORCE | LJ_KICK | LJ_STOREF: return launch_is<double, LJ_FORCE | LJ_KICK | LJ_STOREF>(v, islots, kick_dt, s);
    case LJ_ENERGY: return launch_is<double, LJ_ENERGY>(v, islots, kick_dt, s);
    case LJ_ENERGY | LJ_FORCE | LJ_STOREF: return launch_is<double, LJ_ENERGY | LJ_FORCE | LJ_STOREF>(v, islots, kick_dt, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace chmc
