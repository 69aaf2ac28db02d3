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
// over list slots: the draw of the slot's atom (the same Philox key) goes to pl in slot order.  With
// spec the chains are copies of chain 0 running its next trajectories: copy k draws the momenta of
// trajectory (counter + 1 + k) with chain 0's key.
__global__ __launch_bounds__(256) void k_heatbath(DevView v, uint64_t seed, long long traj_explicit, int count, int slots,
                                                  int spec)
{
    const Chunk ck = v.chunks[blockIdx.x];
    const int ch = ck.chain;
    const ChainDev& C = v.chains[ch];
    const uint64_t traj = traj_explicit >= 0 ? (uint64_t)traj_explicit : (uint64_t)(*v.traj + 1 + (spec ? ch : 0));
    const uint64_t key_chain = spec ? 0 : (uint64_t)ch;
    const double sd = sqrt(C.mass / C.beta);
    const double twopi = 6.283185307179586;
    const int N = v.Ntot;
    double s = 0.0;
    const int* lperm = slots ? list_buf(v, *list_cur(v, ch)).lperm : nullptr;
    double* pout = slots ? v.pl : v.p;
    if (slots && ck.begin == C.atom_off && threadIdx.x == 0) *list_pord(v, ch) = -1;   // pl in the current order
    for (int i = ck.begin + (int)threadIdx.x; i < ck.end; i += 256) {
        const int a = slots ? lperm[i] : i;       // i: the atom, or the slot of atom a
        const u64x4 r = philox4x64_10((uint64_t)(a - C.atom_off), traj, key_chain, 0, seed, kKeyHeatbath);
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
__device__ __forceinline__ void slot_save_restore(const DevView& v, int s, int ch, bool save, int list_q)
{
    const ChainDev& C = v.chains[ch];
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

// An accepted trajectory's end positions start the next trajectory, while its chain's lists were built
// at some earlier point: a rebuild request for the next trajectory's start check (list parity q), after
// which that trajectory and every rejected one from the same positions keep the whole skin.
__device__ __forceinline__ void request_start_rebuild(const DevView& v, int q, int ch)
{
    atomicOr(v.lreq + q * (v.nchains + 1) + 1 + ch, 1);
    atomicOr(v.lreq + q * (v.nchains + 1), 1);
}

__global__ __launch_bounds__(256) void k_save_restore_slots(DevView v, int mode, int list_q)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 2 && s == 0) *v.traj = *v.traj + 1;
    if (s >= v.Ntot) return;
    const int ch = v.atom_chain[s];               // a chain's slots span the same range as its atoms
    const bool save = mode == 2 ? v.state[ch].accepted != 0 : mode == 0;
    slot_save_restore(v, s, ch, save, list_q);
    if (mode == 2 && save && v.start_rebuild && s == v.chains[ch].atom_off) request_start_rebuild(v, list_q, ch);
}

// ---- speculative trajectories: the context's chains are K identical copies of one chain ----------
// Copy k runs trajectory t0 + 1 + k (t0 = the device counter) from the chain's state after t0.  The
// sequential chain would run trajectory t0 + 1 + k from that same state exactly when trajectories
// t0 + 1 .. t0 + k were all rejected, so the copies are tested in order with the Metropolis step of
// k_metropolis against the state after t0; the first accepted copy closes the batch (the copies after
// it started from a state the chain no longer holds and are discarded), and so does the last trajectory
// of the host's window, spec_ctl[3] (a batch after it consumes nothing).  Every workgroup evaluates the same decision from the same inputs; workgroup 0
// writes the records, the control words and the chain state after the batch, which the broadcast
// commits (the kernel that reads the counter and the state never writes them).
//   spec_ctl[0]: copy whose end state the chain takes (-1: none accepted; every copy restores the
//                start state), [1]: trajectories the batch consumed, [2]: copy of the last consumed
//                trajectory (its post-MD momenta are the chain's at the end of a run), [3]: last
//                trajectory of the window (set by the host before the window's batches, so that graph
//                replays captured in an earlier window read it too)
struct SpecDecision {
    int j, consumed;
    ChainState st;
};

__device__ SpecDecision spec_decide(const DevView& v, const double* __restrict__ uniforms, int therm, int ufirst, int tend,
                                    HmcRecord* __restrict__ rec, int rec_base, int surr)
{
    SpecDecision d;
    d.j = -1;
    d.consumed = 0;
    ChainState st = v.state[0];
    const int t0 = *v.traj;
    for (int k = 0; k < v.nchains && t0 + 1 + k <= tend; k++) {
        const int i = t0 + 1 + k;
        const ChainDev& C = v.chains[k];
        const double* r = &v.red[k * RED_NSLOT];
        const double Ki = r[RED_KI], Kf = r[RED_K], Vf = r[RED_V], Wf = r[RED_W];
        const double dh = C.beta * (Kf + Vf - Ki - st.Vi);
        int acc;
        if (i < therm) {
            acc = 1;
        } else {
            const double u = uniforms[i - ufirst];
            acc = (u < exp(-dh)) ? 1 : 0;
            st.nacc += acc;
        }
        if (acc) { st.Vi = Vf; st.Wi = Wf; }
        st.Ki = Ki;
        st.accepted = acc;
        d.consumed = k + 1;
        if (rec) {
            HmcRecord h;
            h.dh = dh; h.u = st.Vi; h.k = acc ? Kf : Ki; h.w = st.Wi; h.acc = acc; h.surr = surr;
            rec[(size_t)(i - 1 - rec_base) * v.nchains] = h;
        }
        if (acc) {
            d.j = k;
            break;
        }
    }
    d.st = st;
    return d;
}

// the decision (the per-chain sums in red[] are those of k_reduce over the energy's tile partials and
// the heat bath's chunk partials), then per slot: the accepted copy keeps its end state as its start
// state, every other copy restores its start state
__global__ __launch_bounds__(256) void k_spec_restore(DevView v, const double* __restrict__ uniforms, int therm, int ufirst,
                                                      HmcRecord* __restrict__ rec, int rec_base, int surr, int list_q)
{
    __shared__ int sj;
    if (threadIdx.x == 0) {
        const bool lead = blockIdx.x == 0;
        const SpecDecision d = spec_decide(v, uniforms, therm, ufirst, v.spec_ctl[3], lead ? rec : nullptr, rec_base, surr);
        sj = d.j;
        if (lead) {
            v.spec_ctl[0] = d.j;
            v.spec_ctl[1] = d.consumed;
            if (d.consumed > 0) v.spec_ctl[2] = d.consumed - 1;
            *v.spec_st = d.st;
        }
    }
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.Ntot) return;
    const int ch = v.atom_chain[s];
    slot_save_restore(v, s, ch, ch == sj, list_q);
    if (ch == sj && v.start_rebuild && s == v.chains[ch].atom_off)
        request_start_rebuild(v, list_q, ch);            // copied to every copy by the broadcast
}

// commit: advance the counter by the consumed trajectories and give every copy the chain state after
// the batch; then, when a copy j was accepted (or src >= 0 names one), copy its state into every other
// copy: positions and start positions by slot and its current list buffer (rows, counts, slot <-> atom
// maps, rebuild references, FP32 coordinates of both parities, slot cells, cell offsets and segment
// starts, rebuild requests), with slot and atom numbers moved to the target copy's range, so that every
// copy continues exactly as
// copy j would.  The momenta are not copied: the next heat bath draws them.
__global__ __launch_bounds__(256) void k_spec_broadcast(DevView v, int src, int commit, int rows)
{
    if (commit && blockIdx.x == 0 && threadIdx.x == 0) {
        *v.traj += v.spec_ctl[1];
        const ChainState st = *v.spec_st;
        for (int k = 0; k < v.nchains; k++) v.state[k] = st;
    }
    const int j = src >= 0 ? src : v.spec_ctl[0];
    if (j < 0) return;
    const ChainDev& Cj = v.chains[j];
    const int b = *list_cur(v, j);
    const ListBuf L = list_buf(v, b);
    const int Nt = v.Ntot;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    int* __restrict__ lreq = v.lreq;
    for (size_t s = gid; s < (size_t)Nt; s += stride) {
        const int k = v.atom_chain[s];
        if (k == j) continue;
        const ChainDev& Ck = v.chains[k];
        const int r = (int)s - Ck.atom_off, sj = Cj.atom_off + r, delta = Ck.atom_off - Cj.atom_off;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            v.xl[(size_t)d * Nt + s] = v.xl[(size_t)d * Nt + sj];
            v.xl_saved[(size_t)d * Nt + s] = v.xl_saved[(size_t)d * Nt + sj];
        }
        L.nbn[s] = L.nbn[sj];
        L.lperm[s] = L.lperm[sj] + delta;
        L.linv[s] = L.linv[sj] + delta;          // atom Ck.atom_off + r: a chain's atoms and slots share a range
        L.lref[s] = L.lref[sj];
        if (v.entry16) L.lcs[s] = L.lcs[sj];     // cells relative to the chain's first
        L.xsl[s] = L.xsl[sj];
        L.xsl[(size_t)Nt + s] = L.xsl[(size_t)Nt + sj];
        tile_cur_mark(v, Ck, (int)s, b);
        if (r == 0) {
            *list_cur(v, k) = b;
            *list_start(v, k) = b;
            *list_pord(v, k) = -1;
            for (int q = 0; q < 2; q++) lreq[q * (v.nchains + 1) + 1 + k] = lreq[q * (v.nchains + 1) + 1 + j];
        }
    }
    // list rows, 16 bytes per thread (four 32-bit entries with their slots moved by the copies' offset,
    // or eight 16-bit ones, relative to the segment starts, as they are), up to the padded span or the
    // count; without rows (every copy rebuilds its lists next) the rows and segment starts stay
    if (!rows) return;
    const bool e16 = v.entry16 != 0;
    const int per = e16 ? 8 : 4, qn = v.nbcap / per, span = v.lspan;
    const size_t rowb = (size_t)v.nbcap * (e16 ? 2 : 4);
    for (size_t u = gid; u < (size_t)Nt * qn; u += stride) {
        const int s = (int)(u / qn), cq = (int)(u % qn);
        const int k = v.atom_chain[s];
        if (k == j) continue;
        const ChainDev& Ck = v.chains[k];
        const int sj = Cj.atom_off + (s - Ck.atom_off);
        if (per * cq >= max(L.nbn[sj], span)) continue;
        int4 e = reinterpret_cast<const int4*>(L.nb + (size_t)sj * rowb)[cq];
        if (!e16) {
            const int shift = (Ck.atom_off - Cj.atom_off) * (1 << kImageBits);
            e.x += shift; e.y += shift; e.z += shift; e.w += shift;
        }
        reinterpret_cast<int4*>(L.nb + (size_t)s * rowb)[cq] = e;
    }
    // cell offsets and (16-bit entries) segment starts, moved to copy k's slot range
    for (size_t c = gid; c < (size_t)v.Cl; c += stride) {
        const int k = v.lcell_chain[c];
        if (k == j) continue;
        const ChainDev& Ck = v.chains[k];
        const int delta = Ck.atom_off - Cj.atom_off;
        const size_t cj = c - Ck.lcell_off + Cj.lcell_off;
        L.loffs[c] = L.loffs[cj] + delta;
        if (!e16) continue;
        const int4* src = reinterpret_cast<const int4*>(L.ctab + cj * kSegStride);
        int4* dst = reinterpret_cast<int4*>(L.ctab + c * kSegStride);
#pragma unroll
        for (int w = 0; w < kSegStride / 4; w++) {
            int4 e = src[w];
            e.x += delta; e.y += delta; e.z += delta; e.w += delta;
            dst[w] = e;
        }
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
        v.p[d * N + a] = lp_round(v.p[d * N + a] - kdt * fd, v.lp_scale);
    }
}

hipError_t launch_heatbath(const DevView& v, uint64_t seed, long long traj_or_neg, bool count, bool slots, hipStream_t s)
{
    hipLaunchKernelGGL(k_heatbath, dim3(v.nchunks), dim3(256), 0, s, v, seed, traj_or_neg, count ? 1 : 0, slots ? 1 : 0, 0);
    return hipGetLastError();
}

hipError_t launch_heatbath_spec(const DevView& v, uint64_t seed, hipStream_t s)
{
    hipLaunchKernelGGL(k_heatbath, dim3(v.nchunks), dim3(256), 0, s, v, seed, -1LL, 0, 1, 1);
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

hipError_t launch_spec_restore(const DevView& v, const double* uniforms, int therm, int ufirst, HmcRecord* rec,
                               int rec_base, int surr, int list_q, hipStream_t s)
{
    hipLaunchKernelGGL(k_spec_restore, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, uniforms, therm, ufirst, rec,
                       rec_base, surr, list_q);
    return hipGetLastError();
}

hipError_t launch_spec_broadcast(const DevView& v, int src, bool commit, bool rows, hipStream_t s)
{
    hipLaunchKernelGGL(k_spec_broadcast, dim3(1024), dim3(256), 0, s, v, src, commit ? 1 : 0, rows ? 1 : 0);
    return hipGetLastError();
}

// end of a speculative run: copy 0 (the chain) takes the atom-order momenta of copy spec_ctl[2], which ran
// the run's last trajectory
__global__ __launch_bounds__(256) void k_spec_final_momenta(DevView v)
{
    const int last = v.spec_ctl[2];
    if (last <= 0 || last >= v.nchains) return;
    const ChainDev& C0 = v.chains[0];
    const ChainDev& Cl = v.chains[last];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C0.natoms) return;
    const size_t Nt = (size_t)v.Ntot;
#pragma unroll
    for (int d = 0; d < 3; d++) v.p[d * Nt + C0.atom_off + i] = v.p[d * Nt + Cl.atom_off + i];
}

hipError_t launch_spec_final_momenta(const DevView& v, hipStream_t s)
{
    const int n = v.Ntot / v.nchains;              // the copies' common size
    hipLaunchKernelGGL(k_spec_final_momenta, dim3((n + 255) / 256), dim3(256), 0, s, v);
    return hipGetLastError();
}

// chain states of a run start: V_i and W_i of the start configuration from the per-chain sums, the rest zero
__global__ void k_state_init(DevView v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= v.nchains) return;
    ChainState st{};
    st.Vi = v.red[(size_t)k * RED_NSLOT + RED_V];
    st.Wi = v.red[(size_t)k * RED_NSLOT + RED_W];
    v.state[k] = st;
}

hipError_t launch_state_init(const DevView& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_state_init, dim3((v.nchains + 255) / 256), dim3(256), 0, s, v);
    return hipGetLastError();
}

hipError_t launch_kick(const DevView& v, double dt, bool fp64, hipStream_t s)
{
    hipLaunchKernelGGL(k_kick, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, dt, fp64 ? 1 : 0);
    return hipGetLastError();
}

}  // namespace chmc
