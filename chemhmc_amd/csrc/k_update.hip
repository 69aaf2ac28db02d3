// Heat bath, Hamiltonian reductions, Metropolis and the accept/restore copy -- the device side of
// HMC_class::run (code/modules/HMC.cpp:61-136), with no host round trip per trajectory.
//   k_heatbath    : p ~ N(0, m/beta) per component (identical_particles.cpp:287-296), Philox4x64-10,
//                   per-chunk partials of K_i, optionally the count phase of the binning of x
//   k_kinetic     : per-chunk partials of K = sum |p|^2 / 2m (identical_particles.cpp:465-468)
//   k_reduce      : one workgroup per chain sums per-cell / per-chunk partials in a fixed order
//   k_metropolis  : the same sums for V_f, W_f, K_f, K_i, then dh = beta (K_f + V_f - K_i - V_i),
//                   thermalisation and u < exp(-dh) (HMC.cpp:88-117), one workgroup per chain
//   k_save_restore: h_x <- x on accept, x <- h_x on reject (HMC.cpp:103,107,111); advances the
//                   device trajectory counter
// Every sum is a fixed-order tree, so results are bitwise reproducible run to run.
#include "chmc_internal.hpp"
#include "bin_device.hpp"
#include "philox.hpp"

#pragma clang fp contract(off)

namespace chmc {

// fixed-order sum of one value per thread of a 256-thread block; result valid in thread 0
__device__ __forceinline__ double block_sum_256(double s)
{
    __shared__ double wpart[4];
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) s += __shfl_xor(s, m, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) wpart[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) t = ((wpart[0] + wpart[1]) + wpart[2]) + wpart[3];
    return t;
}

__global__ __launch_bounds__(256) void k_kinetic(DevView v)
{
    const Chunk ck = v.chunks[blockIdx.x];
    const ChainDev& C = v.chains[ck.chain];
    const int N = v.Ntot;
    double s = 0.0;
    for (int a = ck.begin + (int)threadIdx.x; a < ck.end; a += 256) {
        const double px = v.p[a], py = v.p[N + a], pz = v.p[2 * N + a];
        s += (px * px + py * py + pz * pz) / (2 * C.mass);
    }
    const double t = block_sum_256(s);
    if (threadIdx.x == 0) v.chunk_part[blockIdx.x] = t;
}

// heat bath p ~ N(0, m/beta) (identical_particles.cpp:287-296), K_i partials of the fresh momenta and,
// optionally, the count phase of the binning of the unchanged positions.  With slots the chunk runs
// over list slots: the draw of the slot's atom (the same Philox key) goes to pl in slot order.
__global__ __launch_bounds__(256) void k_heatbath(DevView v, uint64_t seed, long long traj_explicit, int count, int slots)
{
    const Chunk ck = v.chunks[blockIdx.x];
    const int ch = ck.chain;
    const ChainDev& C = v.chains[ch];
    const uint64_t traj = traj_explicit >= 0 ? (uint64_t)traj_explicit : (uint64_t)(*v.traj + 1);
    const double sd = sqrt(C.mass / C.beta);
    const double twopi = 6.283185307179586;
    const int N = v.Ntot;
    double s = 0.0;
    const int* lperm = slots ? list_buf(v, *list_cur(v, ch)).lperm : nullptr;
    double* pout = slots ? v.pl : v.p;
    if (slots && ck.begin == C.atom_off && threadIdx.x == 0) *list_pord(v, ch) = -1;   // pl in the current order
    for (int i = ck.begin + (int)threadIdx.x; i < ck.end; i += 256) {
        const int a = slots ? lperm[i] : i;       // i: the atom, or the slot of atom a
        const u64x4 r = philox4x64_10((uint64_t)(a - C.atom_off), traj, (uint64_t)ch, 0, seed, kKeyHeatbath);
        const double r0 = sqrt(-2.0 * log(u01_oo(r.v[0])));
        const double a0 = twopi * u01_co(r.v[1]);
        const double r1 = sqrt(-2.0 * log(u01_oo(r.v[2])));
        const double a1 = twopi * u01_co(r.v[3]);
        const double px = (r0 * cos(a0)) * sd, py = (r0 * sin(a0)) * sd, pz = (r1 * cos(a1)) * sd;
        pout[i] = px;
        pout[N + i] = py;
        pout[2 * N + i] = pz;
        s += (px * px + py * py + pz * pz) / (2 * C.mass);
        if (count) {
            const double x3[3] = { v.x[a], v.x[N + a], v.x[2 * N + a] };
            count_atom(v, C, a, v.invperm[a], v.cell_of[a], x3);
        }
    }
    const double t = block_sum_256(s);
    if (threadIdx.x == 0) v.chunk_part[blockIdx.x] = t;
}

// per-chain sums (one 256-thread workgroup per chain) of the LJ sweep's per-cell partials and/or the
// per-chunk kinetic partials
__device__ void chain_sums(const DevView& v, int ch, int what)
{
    const ChainDev& C = v.chains[ch];
    double* red = &v.red[ch * RED_NSLOT];
    if (what & SUM_CELLS) {
        double su = 0.0, sw = 0.0, sn = 0.0, sk = 0.0;
        for (int q = threadIdx.x; q < C.ncells; q += 256) {
            const double* cp = &v.cell_part[4 * (C.cell_off + q)];
            su += cp[0]; sw += cp[1]; sn += cp[2]; sk += cp[3];
        }
        su = block_sum_256(su);
        sw = block_sum_256(sw);
        sn = block_sum_256(sn);
        sk = block_sum_256(sk);
        if (threadIdx.x == 0) {
            red[RED_V] = 2 * C.eps * su;      // identical_particles.cpp:392, pairs counted twice
            red[RED_W] = 24 * C.eps * sw;     // sum_{i<j} 48 eps (sr12 - sr6/2)
            red[RED_NINT] = sn;               // ordered interacting pairs
            red[RED_K] = sk;
        }
    }
    if (what & SUM_TILES) {
        double su = 0.0, sw = 0.0, sn = 0.0, sk = 0.0;
        for (int q = threadIdx.x; q < C.ntiles; q += 256) {
            const double* tp = &v.tile_part[4 * (C.tile_first + q)];
            su += tp[0]; sw += tp[1]; sn += tp[2]; sk += tp[3];
        }
        su = block_sum_256(su);
        sw = block_sum_256(sw);
        sn = block_sum_256(sn);
        sk = block_sum_256(sk);
        if (threadIdx.x == 0) {
            red[RED_V] = 2 * C.eps * su;
            red[RED_W] = 24 * C.eps * sw;
            red[RED_NINT] = sn;
            red[RED_K] = sk;
        }
    }
    if (what & (SUM_CHUNKS_KI | SUM_CHUNKS_K)) {
        double q = 0.0;
        for (int k = threadIdx.x; k < C.nchunks; k += 256) q += v.chunk_part[C.chunk_first + k];
        q = block_sum_256(q);
        if (threadIdx.x == 0) red[(what & SUM_CHUNKS_KI) ? RED_KI : RED_K] = q;
    }
}

__global__ __launch_bounds__(256) void k_reduce(DevView v, int what)
{
    chain_sums(v, blockIdx.x, what);
}

// One workgroup per chain: V_f, W_f, K_f from the energy sweep's partials, K_i from the heat bath's,
// then the Metropolis decision.
__global__ __launch_bounds__(256) void k_metropolis(DevView v, const double* __restrict__ uniforms, int ustride,
                                                    int therm, int ufirst, HmcRecord* __restrict__ rec, int rec_base,
                                                    int surr, int what)
{
    const int ch = blockIdx.x;
    chain_sums(v, ch, what | SUM_CHUNKS_KI);
    if (threadIdx.x != 0) return;
    const int i = *v.traj + 1;                 // 1-based index of this trajectory (first_traj = istart + 1)
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
        h.dh = dh; h.u = st.Vi; h.k = acc ? Kf : Ki; h.w = st.Wi; h.acc = acc; h.surr = surr;
        rec[(size_t)(i - 1 - rec_base) * v.nchains + ch] = h;
    }
}

// mode 0: save (x_saved = x), 1: restore (x = x_saved), 2: per chain by the Metropolis decision, which
// also closes the trajectory on the device counter (read by the next heat bath and Metropolis)
// list_q >= 0 (neighbour-list MD at list parity list_q): a save makes the current list buffer the start
// buffer; a restore makes the start buffer current again and recomputes its FP32 coordinates of parity
// list_q for the restored positions (they are within skin / 2 of the start buffer's rebuild positions,
// which it was valid for when they were saved, so no rebuild is needed)
__global__ __launch_bounds__(256) void k_save_restore(DevView v, int mode, int list_q)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 2 && a == 0) *v.traj = *v.traj + 1;
    if (a >= v.Ntot) return;
    const int ch = v.atom_chain[a];
    const ChainDev& C = v.chains[ch];
    int save = mode == 0;
    if (mode == 2) save = v.state[ch].accepted;
    const int N = v.Ntot;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (save) v.x_saved[d * N + a] = v.x[d * N + a];
        else v.x[d * N + a] = v.x_saved[d * N + a];
    }
    if (list_q < 0) return;
    if (save) {
        if (a == C.atom_off) *list_start(v, ch) = *list_cur(v, ch);
        return;
    }
    const int b = *list_start(v, ch);
    if (a == C.atom_off) *list_cur(v, ch) = b;
    tile_cur_mark(v, C, a, b);            // every tile of the chain is marked by one of its threads
    const ListBuf L = list_buf(v, b);
    const int slot = L.linv[a];
    float f3[3];
    const float4 ref = L.lref[slot];
#pragma unroll
    for (int d = 0; d < 3; d++) f3[d] = (float)list_rel(C, ref, d, v.x_saved[d * N + a]);
    L.xsl[(size_t)list_q * N + slot] = make_float4(f3[0], f3[1], f3[2], __int_as_float(a));
}

// The same on the slot-ordered state of list MD: xl_saved <- xl on save (the current list buffer becomes
// the start buffer), xl <- xl_saved on restore (the start buffer becomes current again, and its FP32
// coordinates of parity list_q are recomputed for the restored positions)
__global__ __launch_bounds__(256) void k_save_restore_slots(DevView v, int mode, int list_q)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 2 && s == 0) *v.traj = *v.traj + 1;
    if (s >= v.Ntot) return;
    const int ch = v.atom_chain[s];               // a chain's slots span the same range as its atoms
    const ChainDev& C = v.chains[ch];
    int save = mode == 0;
    if (mode == 2) save = v.state[ch].accepted;
    const int N = v.Ntot;
    if (save) {
#pragma unroll
        for (int d = 0; d < 3; d++) v.xl_saved[d * N + s] = v.xl[d * N + s];
        if (s == C.atom_off) *list_start(v, ch) = *list_cur(v, ch);
        return;
    }
    const int b = *list_start(v, ch);
    if (s == C.atom_off) {
        // the momenta stay in the slot order of the buffer the trajectory ended in (the reference keeps
        // the post-MD momenta of a rejected trajectory, HMC.cpp:111): recorded for the export
        const int cur = *list_cur(v, ch);
        *list_pord(v, ch) = cur != b ? cur : -1;
        *list_cur(v, ch) = b;
    }
    tile_cur_mark(v, C, s, b);
    const ListBuf L = list_buf(v, b);
    float f3[3];
    const float4 ref = L.lref[s];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double xv = v.xl_saved[d * N + s];
        v.xl[d * N + s] = xv;
        f3[d] = (float)list_rel(C, ref, d, xv);
    }
    L.xsl[(size_t)list_q * N + s] = make_float4(f3[0], f3[1], f3[2], 0.f);
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
        v.p[d * N + a] = lp_round(v.p[d * N + a] - kdt * fd, v.lp_scale);
    }
}

hipError_t launch_heatbath(const DevView& v, uint64_t seed, long long traj_or_neg, bool count, bool slots, hipStream_t s)
{
    hipLaunchKernelGGL(k_heatbath, dim3(v.nchunks), dim3(256), 0, s, v, seed, traj_or_neg, count ? 1 : 0, slots ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_kinetic(const DevView& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_kinetic, dim3(v.nchunks), dim3(256), 0, s, v);
    return hipGetLastError();
}

hipError_t launch_reduce(const DevView& v, int what, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce, dim3(v.nchains), dim3(256), 0, s, v, what);
    return hipGetLastError();
}

hipError_t launch_metropolis(const DevView& v, const double* uniforms, int ustride, int therm, int ufirst,
                             HmcRecord* rec, int rec_base, int surr, int what, hipStream_t s)
{
    hipLaunchKernelGGL(k_metropolis, dim3(v.nchains), dim3(256), 0, s, v, uniforms, ustride, therm, ufirst, rec, rec_base,
                       surr, what);
    return hipGetLastError();
}

hipError_t launch_save_restore(const DevView& v, int mode, int list_q, hipStream_t s)
{
    hipLaunchKernelGGL(k_save_restore, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, mode, list_q);
    return hipGetLastError();
}

hipError_t launch_save_restore_slots(const DevView& v, int mode, int list_q, hipStream_t s)
{
    hipLaunchKernelGGL(k_save_restore_slots, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, mode, list_q);
    return hipGetLastError();
}

hipError_t launch_kick(const DevView& v, double dt, bool fp64, hipStream_t s)
{
    hipLaunchKernelGGL(k_kick, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, dt, fp64 ? 1 : 0);
    return hipGetLastError();
}

}  // namespace chmc
