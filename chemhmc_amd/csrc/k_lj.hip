// Lennard-Jones sweep over the 27 neighbour cells: the replacement of Tag_force_binning
// (code/modules/identical_particles.cpp:567-639), Tag_potential_binning (:395-454) and the kick of
// update_momenta (:676-694), with a pair virial the reference does not have.
//
// One workgroup of W wavefronts per i-cell:
//   1. lanes 0..26 of wave 0 resolve the 27 neighbour cells (with the reference's explicit image
//      shift when the neighbour index wraps, :584-600); a prefix sum gives the concatenated j-list;
//   2. all W*64 threads stage the j-list into LDS at once (binary search of the segment, one load
//      each, all in flight together), with the image shift folded into the staged coordinates;
//   3. inside a wave, lanes are IS i-slots x (64/IS) j-groups (lane = g*IS + ii): i-atom ii of the
//      wave's current i-chunk against every (64/IS)-th j entry, one broadcast LDS read per pair;
//      waves take i-chunks round-robin; a butterfly over the j-groups sums the partials.
// FP32 path: coordinates relative to the cell origin (|d| < 2 cell sides), per-atom FP32 sums,
//   +grad U stored and/or the FP64 momenta kicked in place (p -= (dt*beta)*f).
// FP64 path: coordinates relative to the i-cell origin in FP64, reciprocal from v_rcp_f64 with two
//   Newton steps; optional per-atom energy / virial / kinetic sums, reduced deterministically per cell
//   (fixed wave order) into per-cell partials that the chain reduction sums in cell order.
// Self-exclusion is r^2 > 0: the staged copy of atom i is bitwise equal to its own coordinates.
// LJ_DRIFT (FP32 path only, which reads positions through the sorted FP32 copy): right after its kick
// each atom also takes the integrator's next drift and is counted into its new cell (the count phase of
// the next binning).
#include "chmc_internal.hpp"
#include "bin_device.hpp"

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

__device__ __forceinline__ float rcp_(float r2) { return __builtin_amdgcn_rcpf(r2); }
__device__ __forceinline__ double rcp_(double r2)
{
    double y = __builtin_amdgcn_rcp(r2);
    double e = __builtin_fma(-r2, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-r2, y, 1.0);
    return __builtin_fma(y, e, y);
}
__device__ __forceinline__ float  fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename Real, int IS, int kWaves, int MODE>
__global__ __launch_bounds__(64 * kWaves) void k_lj(DevView v, double kick_dt, double drift_dt, int nblocks)
{
    constexpr int JG = 64 / IS;
    constexpr int NT = 64 * kWaves;
    __shared__ Stage<Real> sj;
    __shared__ int seg_beg[27];
    __shared__ int seg_pref[28];
    __shared__ Real seg_sh[27][3];
    __shared__ double wsum[kWaves][4];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int c = xcd_remap(blockIdx.x, nblocks);
    if (c < v.Ctot) {
    const int ch = v.cell_chain[c];
    const ChainDev& C = v.chains[ch];
    const int nx = C.nbin[0], ny = C.nbin[1], nz = C.nbin[2];
    const int lc = c - C.cell_off;
    const int bz = lc / (nx * ny);
    const int by = (lc - bz * nx * ny) / nx;
    const int bx = lc - nx * (by + bz * ny);
    const int* offs = offs_cur(v);
    const int* perm = perm_cur(v);
    const int ibeg = offs[c];
    const int ni = offs[c + 1] - ibeg;
    const int N = v.Ntot;

    double cell_u = 0.0, cell_w = 0.0, cell_n = 0.0, cell_k = 0.0;
    if (ni > 0) {
        // ---- 1. neighbour segments -----------------------------------------------------------
        if (wave == 0) {
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
                const int o0 = offs[jc];
                seg_beg[lane] = o0;
                cnt = offs[jc + 1] - o0;
            }
            int incl = cnt;
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                int t = __shfl_up(incl, o, 64);
                if (lane >= o) incl += t;
            }
            if (lane < 27) seg_pref[lane + 1] = incl;
            if (lane == 0) seg_pref[0] = 0;
        }
        __syncthreads();
        const int nj = seg_pref[27];

        // ---- 2. staging of j-list entries [jlo, jhi): every thread, binary search of its segment ----
        auto stage = [&](int jlo, int jhi) {
            for (int t = jlo + tid; t < jhi; t += NT) {
                int k = 0;
#pragma unroll
                for (int step = 16; step > 0; step >>= 1)
                    if (k + step < 27 && seg_pref[k + step] <= t) k += step;
                const int src = seg_beg[k] + (t - seg_pref[k]);
                if constexpr (sizeof(Real) == 4) {
                    const float4 q = v.xs[src];
                    sj.q[t - jlo] = make_float4(q.x + seg_sh[k][0], q.y + seg_sh[k][1], q.z + seg_sh[k][2], 0.f);
                } else {
                    const int a = perm[src];
                    sj.x[t - jlo] = v.x[a] + seg_sh[k][0];
                    sj.y[t - jlo] = v.x[N + a] + seg_sh[k][1];
                    sj.z[t - jlo] = v.x[2 * N + a] + seg_sh[k][2];
                }
            }
        };
        const bool single = nj <= kMaxJ;
        if (single) {
            stage(0, nj);
            __syncthreads();
        }

        // ---- 3. i-chunks (round-robin over waves) x j-groups ----------------------------------------
        const int ii = lane % IS, g = lane / IS;
        const Real rc2 = (sizeof(Real) == 4) ? (Real)C.rc2_f : (Real)C.rc2;
        const Real sig2 = (sizeof(Real) == 4) ? (Real)C.sigma2_f : (Real)C.sigma2;
        const int nchunk = (ni + IS - 1) / IS;
        // every wave runs the same number of chunk iterations so that the multi-pass staging barriers match
        const int rounds = (nchunk + kWaves - 1) / kWaves;
        for (int rnd = 0; rnd < rounds; rnd++) {
            const int chunk = rnd * kWaves + wave;
            const int i = chunk * IS + ii;
            const bool ival = i < ni;
            const int ic = ival ? i : ni - 1;
            Real xi, yi, zi;
            int atom;
            if constexpr (sizeof(Real) == 4) {
                const float4 q = v.xs[ibeg + ic];
                xi = q.x; yi = q.y; zi = q.z;
                atom = __float_as_int(q.w);
            } else {
                atom = perm[ibeg + ic];
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
                const int jn = (chunk < nchunk) ? jhi - jlo : 0;
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
            if (chunk >= nchunk) continue;   // wave-uniform
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
                    double pn[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        if constexpr ((MODE & LJ_STOREF) != 0) {
                            if constexpr (sizeof(Real) == 4) v.f32[d * N + atom] = f3[d];
                            else v.f64[d * N + atom] = f3[d];
                        }
                        if constexpr ((MODE & LJ_KICK) != 0) {
                            const double fd = (double)f3[d];
                            pn[d] = lp_round(v.p[d * N + atom] - kdt * fd, v.lp_scale);
                            v.p[d * N + atom] = pn[d];
                        }
                    }
                    if constexpr ((MODE & LJ_DRIFT) != 0) {
                        // the following update_positions (identical_particles.cpp:651-669) of this atom,
                        // then the count phase of the next binning at its new position
                        const double coeff_x = C.coeff_x;
                        double x3[3];
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            double xv = v.x[d * N + atom];
                            xv += drift_dt * coeff_x * pn[d];
                            xv = lp_round(xv, v.lp_scale);
                            xv = wrap_box(xv, C.L[d]);
                            v.x[d * N + atom] = xv;
                            x3[d] = xv;
                        }
                        count_atom(v, C, atom, ibeg + i, c, x3);
                    }
                }
            }
            if constexpr ((MODE & LJ_ENERGY) != 0) {
                double eu = (g == 0 && ival) ? (double)au : 0.0;
                double ew = (g == 0 && ival) ? (double)aw : 0.0;
                int en = (g == 0 && ival) ? an : 0;
                // kinetic energy of the i-atom: every atom is an i-atom exactly once (K_f, HMC.cpp:90)
                double ek = 0.0;
                if (g == 0 && ival) {
                    const double px = v.p[atom], py = v.p[N + atom], pz = v.p[2 * N + atom];
                    ek = (px * px + py * py + pz * pz) / (2 * C.mass);
                }
#pragma unroll
                for (int m = 1; m < IS; m <<= 1) {
                    eu += __shfl_xor(eu, m, 64);
                    ew += __shfl_xor(ew, m, 64);
                    en += __shfl_xor(en, m, 64);
                    ek += __shfl_xor(ek, m, 64);
                }
                cell_u += eu;
                cell_w += ew;
                cell_n += (double)en;
                cell_k += ek;
            }
        }
    }

    if constexpr ((MODE & LJ_ENERGY) != 0) {
        // per-cell partials in fixed wave order, then the last workgroup of the chain sums them in cell order
        if (lane == 0) {
            wsum[wave][0] = cell_u;
            wsum[wave][1] = cell_w;
            wsum[wave][2] = cell_n;
            wsum[wave][3] = cell_k;
        }
        __syncthreads();
        if (tid == 0) {
            double su = 0.0, sw = 0.0, sn = 0.0, sk = 0.0;
            for (int w = 0; w < kWaves; w++) { su += wsum[w][0]; sw += wsum[w][1]; sn += wsum[w][2]; sk += wsum[w][3]; }
            v.cell_part[4 * c] = su;
            v.cell_part[4 * c + 1] = sw;
            v.cell_part[4 * c + 2] = sn;
            v.cell_part[4 * c + 3] = sk;
        }
    }
    }   // c < Ctot
}

template <typename Real, int IS, int W, int MODE>
static void launch_one(const DevView& v, double kick_dt, double drift_dt, hipStream_t s)
{
    const int nblocks = ((v.Ctot + 7) / 8) * 8;
    hipLaunchKernelGGL((k_lj<Real, IS, W, MODE>), dim3(nblocks), dim3(64 * W), 0, s, v, kick_dt, drift_dt, nblocks);
}

template <typename Real, int MODE>
static hipError_t launch_tiled(const DevView& v, int tiling, double kick_dt, double drift_dt, hipStream_t s)
{
    switch (tiling) {
    case 41: launch_one<Real, 4, 1, MODE>(v, kick_dt, drift_dt, s); break;
    case 42: launch_one<Real, 4, 2, MODE>(v, kick_dt, drift_dt, s); break;
    case 44: launch_one<Real, 4, 4, MODE>(v, kick_dt, drift_dt, s); break;
    case 81: launch_one<Real, 8, 1, MODE>(v, kick_dt, drift_dt, s); break;
    case 84: launch_one<Real, 8, 4, MODE>(v, kick_dt, drift_dt, s); break;
    case 161: launch_one<Real, 16, 1, MODE>(v, kick_dt, drift_dt, s); break;
    case 162: launch_one<Real, 16, 2, MODE>(v, kick_dt, drift_dt, s); break;
    default: launch_one<Real, 8, 2, MODE>(v, kick_dt, drift_dt, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_lj(const DevView& v, int mode, bool fp64, int tiling, double kick_dt, double drift_dt, hipStream_t s)
{
    constexpr int FK = LJ_FORCE | LJ_KICK;
    if (!fp64) {
        switch (mode) {
        case FK: return launch_tiled<float, FK>(v, tiling, kick_dt, drift_dt, s);
        case LJ_FORCE | LJ_STOREF: return launch_tiled<float, LJ_FORCE | LJ_STOREF>(v, tiling, kick_dt, drift_dt, s);
        case FK | LJ_STOREF: return launch_tiled<float, FK | LJ_STOREF>(v, tiling, kick_dt, drift_dt, s);
        case FK | LJ_DRIFT: return launch_tiled<float, FK | LJ_DRIFT>(v, tiling, kick_dt, drift_dt, s);
        case FK | LJ_STOREF | LJ_DRIFT: return launch_tiled<float, FK | LJ_STOREF | LJ_DRIFT>(v, tiling, kick_dt, drift_dt, s);
        default: return hipErrorInvalidValue;
        }
    }
    // the FP64 sweeps run once per trajectory or in FP64 MD: one fixed tiling
    switch (mode) {
    case FK: launch_one<double, 8, 2, FK>(v, kick_dt, drift_dt, s); break;
    case LJ_FORCE | LJ_STOREF: launch_one<double, 8, 2, LJ_FORCE | LJ_STOREF>(v, kick_dt, drift_dt, s); break;
    case FK | LJ_STOREF: launch_one<double, 8, 2, FK | LJ_STOREF>(v, kick_dt, drift_dt, s); break;
    case LJ_ENERGY: launch_one<double, 8, 2, LJ_ENERGY>(v, kick_dt, drift_dt, s); break;
    case LJ_ENERGY | LJ_FORCE | LJ_STOREF: launch_one<double, 8, 2, LJ_ENERGY | LJ_FORCE | LJ_STOREF>(v, kick_dt, drift_dt, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace chmc
