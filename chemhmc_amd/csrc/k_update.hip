// Heat bath, Hamiltonian reductions, Metropolis and the accept/restore copy -- the device side of
// HMC_class::run (code/modules/HMC.cpp:61-136), with no host round trip per trajectory.
//   k_heatbath   : p ~ N(0, m/beta) per component (identical_particles.cpp:287-296), Philox4x64-10
//   k_kinetic    : K = sum |p|^2 / 2m (identical_particles.cpp:465-468), two-level deterministic sum
//   k_metropolis : dh = beta (K_f + V_f - K_i - V_i), thermalisation, u < exp(-dh) (HMC.cpp:88-117)
//   k_save_restore: h_x <- x on accept, x <- h_x on reject (HMC.cpp:103,107,111)
#include "chmc_internal.hpp"
#include "philox.hpp"

#pragma clang fp contract(off)

namespace chmc {

__global__ __launch_bounds__(256) void k_heatbath(DevView v, uint64_t seed, long long traj_explicit)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= v.Ntot) return;
    const int ch = v.atom_chain[a];
    const ChainDev& C = v.chains[ch];
    const uint64_t traj = traj_explicit >= 0 ? (uint64_t)traj_explicit : (uint64_t)(*v.traj + 1);
    const u64x4 r = philox4x64_10((uint64_t)(a - C.atom_off), traj, (uint64_t)ch, 0, seed, kKeyHeatbath);
    const double sd = sqrt(C.mass / C.beta);
    const double twopi = 6.283185307179586;
    const double r0 = sqrt(-2.0 * log(u01_oo(r.v[0])));
    const double a0 = twopi * u01_co(r.v[1]);
    const double r1 = sqrt(-2.0 * log(u01_oo(r.v[2])));
    const double a1 = twopi * u01_co(r.v[3]);
    const int N = v.Ntot;
    v.p[a] = (r0 * cos(a0)) * sd;
    v.p[N + a] = (r0 * sin(a0)) * sd;
    v.p[2 * N + a] = (r1 * cos(a1)) * sd;
}

// fixed-order sum of one value per thread of a 256-thread block; result valid in thread 0
__device__ __forceinline__ double block_sum_256(double s)
{
    __shared__ double wpart[4];
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) s += __shfl_xor(s, m, 64);
    if ((threadIdx.x & 63) == 0) wpart[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) t = ((wpart[0] + wpart[1]) + wpart[2]) + wpart[3];
    return t;
}

__global__ __launch_bounds__(256) void k_kinetic(DevView v, int slot)
{
    __shared__ int last_block;
    const Chunk ck = v.chunks[blockIdx.x];
    const ChainDev& C = v.chains[ck.chain];
    const int N = v.Ntot;
    double s = 0.0;
    for (int a = ck.begin + (int)threadIdx.x; a < ck.end; a += 256) {
        const double px = v.p[a], py = v.p[N + a], pz = v.p[2 * N + a];
        s += (px * px + py * py + pz * pz) / (2 * C.mass);
    }
    const double tot = block_sum_256(s);
    if (threadIdx.x == 0) v.chunk_part[blockIdx.x] = tot;
    __threadfence();
    if (threadIdx.x == 0) last_block = (atomicAdd(&v.ticket[ck.chain], 1) == C.nchunks - 1) ? 1 : 0;
    __syncthreads();
    if (last_block) {
        __threadfence();
        double q = 0.0;
        for (int k = threadIdx.x; k < C.nchunks; k += 256)
            q += __hip_atomic_load(&v.chunk_part[C.chunk_first + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const double K = block_sum_256(q);
        if (threadIdx.x == 0) {
            v.red[ck.chain * RED_NSLOT + slot] = K;
            v.ticket[ck.chain] = 0;
        }
    }
}

// One workgroup walks all chains, then advances the device trajectory counter.
__global__ __launch_bounds__(1024) void k_metropolis(DevView v, const double* __restrict__ uniforms, int ustride,
                                                     int therm, int ufirst, HmcRecord* __restrict__ rec,
                                                     int rec_base)
{
    const int i = *v.traj + 1;                 // 1-based index of this trajectory (first_traj = istart + 1)
    for (int ch = threadIdx.x; ch < v.nchains; ch += blockDim.x) {
        const ChainDev& C = v.chains[ch];
        ChainState st = v.state[ch];
        const double* r = &v.red[ch * RED_NSLOT];
        const double Ki = r[RED_KI], Kf = r[RED_K], Vf = r[RED_V], Wf = r[RED_W];
        const double dh = C.beta * (Kf + Vf - Ki - st.Vi);
        int acc;
        if (i < therm) {
            acc = 1;
        } else {
            // uniforms are consumed in order by the post-thermalisation trajectories of this run
            const double u = uniforms[(size_t)ch * ustride + (i - ufirst)];
            acc = (u < exp(-dh)) ? 1 : 0;
            st.nacc += acc;
        }
        if (acc) { st.Vi = Vf; st.Wi = Wf; }
        st.Ki = Ki;
        st.accepted = acc;
        v.state[ch] = st;
        if (rec) {
            HmcRecord h;
            h.dh = dh; h.u = st.Vi; h.k = Ki; h.w = st.Wi; h.acc = acc; h.pad_ = 0;
            rec[(size_t)(i - 1 - rec_base) * v.nchains + ch] = h;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *v.traj = i;
}

// mode 0: save (x_saved = x), 1: restore (x = x_saved), 2: per chain by the Metropolis decision
__global__ __launch_bounds__(256) void k_save_restore(DevView v, int mode)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= v.Ntot) return;
    int save = mode == 0;
    if (mode == 2) save = v.state[v.atom_chain[a]].accepted;
    const int N = v.Ntot;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (save) v.x_saved[d * N + a] = v.x[d * N + a];
        else v.x[d * N + a] = v.x_saved[d * N + a];
    }
}

// Standalone kick with the stored force: p -= (dt*beta) f  (identical_particles.cpp:686-688)
__global__ __launch_bounds__(256) void k_kick(DevView v, double dt, int fp64)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= v.Ntot) return;
    const ChainDev& C = v.chains[v.atom_chain[a]];
    const int N = v.Ntot;
    const double kdt = dt * C.beta;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double fd = fp64 ? v.f64[d * N + a] : (double)v.f32[d * N + a];
        v.p[d * N + a] = v.p[d * N + a] - kdt * fd;
    }
}

hipError_t launch_heatbath(const DevView& v, uint64_t seed, long long traj_or_neg, hipStream_t s)
{
    hipLaunchKernelGGL(k_heatbath, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, seed, traj_or_neg);
    return hipGetLastError();
}

hipError_t launch_kinetic(const DevView& v, int slot, hipStream_t s)
{
    hipLaunchKernelGGL(k_kinetic, dim3(v.nchunks), dim3(256), 0, s, v, slot);
    return hipGetLastError();
}

hipError_t launch_metropolis(const DevView& v, const double* uniforms, int ustride, int therm, int ufirst,
                             HmcRecord* rec, int rec_base, hipStream_t s)
{
    hipLaunchKernelGGL(k_metropolis, dim3(1), dim3(1024), 0, s, v, uniforms, ustride, therm, ufirst, rec, rec_base);
    return hipGetLastError();
}

hipError_t launch_save_restore(const DevView& v, int mode, hipStream_t s)
{
    hipLaunchKernelGGL(k_save_restore, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, mode);
    return hipGetLastError();
}

hipError_t launch_kick(const DevView& v, double dt, bool fp64, hipStream_t s)
{
    hipLaunchKernelGGL(k_kick, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, dt, fp64 ? 1 : 0);
    return hipGetLastError();
}

}  // namespace chmc
