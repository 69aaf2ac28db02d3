// Internal declarations shared by the HIP translation units of libchmc.
//
// Data layout in HBM (one context = a batch of independent chains, concatenated):
//   x, p, x_saved : double SoA [3][Ntot]      (component d of atom a at d*Ntot + a)
//   f32           : float  SoA [3][Ntot]      (+grad U, FP32 pair sums)
//   f64           : double SoA [3][Ntot]      (+grad U, FP64 path)
//   atom_chain    : int [Ntot]
//   cell_of, slot : int [Ntot]                (cell of each atom at the last binning, arrival slot)
//   newcell       : int [Ntot]                (new cell id, indexed by the atom's previous sorted slot)
//   count         : int [Ctot]                (zero between binnings; the scan clears it)
//   offs_buf      : int [2][Ctot + 1]         (global exclusive scans of the counts: current / next)
//   skey, satom   : int [Ntot]                (unsorted within-cell scatter, sort key and atom)
//   perm_buf      : int [2][Ntot]             (sorted slot -> atom: current / next)
//   invperm       : int [Ntot]                (atom -> current sorted slot)
//   xs            : float4 [Ntot]             (sorted cell-relative FP32 coordinates, atom id in w)
//   bflag         : int [2]                   (locality-violation flag, by target buffer)
//   par (argument): which perm/offs buffer is current; the host flips it after every placement
//   cell_chain    : int [Ctot]
// Chains are contiguous in both the atom and the cell index spaces, so the global scan of the
// per-cell counts equals chain atom offset + the chain's own exclusive scan.
//
// Neighbour-list MD (FP32 forces between rebuilds; cells of side >= rc + skin on a second grid):
//   lcount        : int [2][Cl]               (member counts; the host alternates the buffer with every
//                                              rebuild launch, which clears the other one)
//   loffs         : int [Cl + 1]              (exclusive scan: first compact slot of each list cell)
//   lmem          : float4 [sum_chains lncells * lcap]  (members of each cell in arrival order: FP32
//                                              position relative to the cell origin, atom id in w)
//   lmemx, lmemp  : double [3][same]          (the members' FP64 positions and momenta)
//   xl, pl        : double SoA [3][Ntot] by slot  (positions and momenta in the current buffer's slot
//                                              order while list MD owns the state: the sweeps read
//                                              and write them coalesced; imported from / exported to
//                                              x and p (atom order) at the boundaries of an MD run)
//   xl_saved      : double SoA [3][Ntot] by slot  (h_x of the device HMC loop, in the slot order of
//                                              the trajectory's start buffer)
//   lperm / linv  : int [Ntot]                (compact slot -> atom / atom -> slot, atom-id order
//                                              inside a cell)
//   lref          : float4 [Ntot] by slot     (rebuild reference: the FP32 position relative to the
//                                              rebuild cell origin, r0, and the cell's grid coordinates
//                                              bx | by << 10 | bz << 20 in w)
//   xsl           : float4 [2][Ntot] by slot  (current position relative to the rebuild cell origin,
//                                              FP32, atom id in w; double-buffered across drifts)
//   nb            : [Ntot][nbcap] list entries in one of two formats (entry16, chosen per context):
//                   32-bit: j slot << 6 | image code (2 bits per dimension, the neighbour cell offset + 1,
//                           bits 0-1 x, 2-3 y, 4-5 z; code kFarCode: no-op padding)
//                   16-bit: segment k << 8 | rank r (the neighbour is member r, in slot order, of
//                           segment k of the slot's list cell, k = 0..26 the 27 neighbour cells at
//                           offsets (k % 3 - 1, k / 3 % 3 - 1, k / 9 - 1); k = kPadSeg: no-op padding)
//   nbn           : int [Ntot]
//   lcs           : int [Ntot] by slot        (16-bit entries: the slot's list cell, relative to its
//                                              chain's first)
//   ctab          : int [Cl][kSegStride]      (16-bit entries: first slot of each segment of each list
//                                              cell, the padding segments holding the cell's own first
//                                              slot, so that neighbour slot = ctab[cell][k] + r)
//   lreq          : int [2][1 + nchains]      (parity q: [0] any chain, [1 + k] chain k needs a
//                                              rebuild before its next sweep; chains rebuild
//                                              independently)
//   lflag         : int [4]                   ([2]: capacity overflow, [3]: chain rebuilds run)
//   The per-slot list state (loffs, lperm, linv, lref, xsl, nb, nbn, lcs, ctab) exists twice: a rebuild writes
//   the buffer that is not the trajectory's start buffer, so a rejected trajectory goes back to the
//   start buffer (valid for the restored positions) instead of rebuilding; the momenta pl are left in the
//   slot order of the buffer the trajectory ended in (lbuf[2 nchains + k], -1 when that is the current
//   one), which the export follows.  lbuf[k] / lbuf[nchains + k]
//   are chain k's current and start buffers; tcur[t] repeats the current buffer of tile t's chain, so
//   that a sweep workgroup finds its buffer with one load that depends on nothing but its index.
//   lscan_tmp     : int [Cl]                  (global exclusive scan of the counts, rebuild scratch)
//   tiles         : 32-slot ranges that never straddle chains (sweep / energy workgroups)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace chmc {

struct ChainDev {
    int atom_off, natoms, cell_off, ncells;
    int nbin[3];
    int chunk_first, nchunks;      // deterministic reduction chunks of this chain
    int pad_;
    int lnbin[3], lcell_off, lncells, lcap, ldense_off, tile_first, ntiles;
    int lpad_;
    double L[3], sizebin[3];
    double rc, rc2, sigma, sigma2, eps, mass, beta;
    float  sizebin_f[3], rc2_f, sigma2_f, eps48_f;
    double lsize[3], half_skin2;   // list cell side, (skin / 2)^2
    float  lsize_f[3], rl2_f;      // list radius^2 = (rc + skin)^2
    double p2max;                  // |p|^2 bound of the lists' check interval: (10 sqrt(m / beta))^2
    double coeff_x;                // beta / m, the drift coefficient (compute_coeff_position)
    double half_skin, vmax;        // sqrt(half_skin2); coeff_x * sqrt(p2max), the largest drift speed
};

// per-chain reduction slots
enum RedSlot { RED_K = 0, RED_V = 1, RED_W = 2, RED_KI = 3, RED_NINT = 4, RED_NSLOT = 5 };

// persistent per-chain HMC state on the device
struct ChainState {
    double Vi, Wi, Ki;
    int nacc, accepted;
    int pad_[2];
};

struct HmcRecord {
    double dh, u, k, w;
    int acc, surr;
};

struct Chunk { int chain, begin, end, pad_; };

constexpr int kMaxJ = 768;         // j-list entries staged in LDS per pass
constexpr int kRedChunk = 256;     // atoms per deterministic reduction chunk
constexpr int kTile = 32;          // slots per neighbour-list sweep workgroup (8 lanes per slot)
// leading entries of every list row that the sweep reads unconditionally: the build pads a row past its
// count up to this span with entries that contribute nothing (32-bit: the slot itself with the image
// code kFarCode, which no neighbour has; 16-bit: segment kPadSeg; the shift tables place either beyond
// any cut-off)
constexpr int kListSpan = 128;
constexpr int kFarCode = 63;
constexpr int kSegStride = 32;     // ctab words per list cell: 27 neighbour segments + padding segments
constexpr int kPadSeg = 27;
constexpr int kRankBits = 8;       // a list cell holds at most 256 members
constexpr uint16_t kPadEntry = (uint16_t)(kPadSeg << kRankBits);
// the unconditional first batch of each list-sweep shape (lanes per atom x entries per lane): 4 lanes
// x 16, 8 x 14, 16 x 7, 8 x 10, 11, 12 or 16 ("810" ... "816"); rows are padded at least this far
__host__ __device__ constexpr int sweep_span(int shape)
{
    return shape == 4 ? 64 : shape == 8 ? 112 : shape == 16 ? 112 : shape == 810 ? 80 : shape == 811 ? 88
         : shape == 812 ? 96 : 128;
}
constexpr float kFarShift = 1.0e18f;
constexpr int kScanCells = 3072;   // list cells of a chain whose counts one build workgroup scans in LDS
constexpr int kScanChainCells = 512;   // batches of chains this large scan per chain in a kernel of its own

// ---- launch wrappers (each defined in the TU that owns the kernel) ----------------------
struct DevView {
    int Ntot, Ctot, nchains, nchunks;
    const ChainDev* chains;
    const int* atom_chain;
    const int* cell_chain;
    const Chunk* chunks;
    double *x, *p, *x_saved, *f64;
    float* f32;
    int *cell_of, *slot, *newcell, *count, *offs_buf, *skey, *satom, *perm_buf, *invperm, *bflag;
    int par;              // current binning buffer (0/1)
    float4* xs;
    double* red;          // [nchains][RED_NSLOT]
    double* chunk_part;   // [nchunks]: per-chunk kinetic energy
    double* cell_part;    // [Ctot * 4]: energy, virial, interacting pairs, kinetic energy
    ChainState* state;    // [nchains]
    int* traj;            // [1] device trajectory counter (1-based index of the last finished one)
    // neighbour-list MD
    int Cl, ntiles, nbcap, lanes;   // lanes: list-sweep shape (4, 8, 16 lanes per slot; 810 ... 816: 8 lanes
                                    // with first batches of 10 ... 16 entries)
    const int* lcell_chain;
    const Chunk* tiles;
    int *lcount, *loffs, *lperm, *linv, *nbn, *lflag, *lreq, *lscan_tmp, *lcs, *ctab;
    int* lrank;           // [dense] rank of each counted member in its list cell (k_lscan_chains)
    char* nb;             // list rows (entry bytes: 2 with entry16, else 4)
    size_t nb_stride;     // bytes of one buffer's rows: Ntot * nbcap * entry bytes
    int entry16;          // 16-bit segment-relative list entries (else 32-bit slot entries)
    float4* lmem;
    double *lmemx, *lmemp; // [3][dense]
    double *xl, *pl, *xl_saved;
    int ldense;           // member slots of all chains (sum of lncells * lcap)
    int lscan_chains;     // a batch of chains of kScanChainCells list cells or more: a scan workgroup per chain
    int lscan_in_build;   // every chain has at most kScanCells list cells: the build scans the counts
    int lcap_max;         // largest list-cell member capacity over the chains (at most 256)
    int* lbuf;            // [3][nchains]: current list buffer, buffer of the trajectory's start, buffer whose
                          // slot order pl is in when that is not the current one (-1 otherwise)
    int* tcur;            // [ntiles]: current list buffer of each tile's chain (kept equal to lbuf)
    float4* lref;
    double* rel64;        // [Ntot][4] by slot: FP64 position relative to the rebuild cell origin (x, y, z, -)
    float4* xsl;
    double* tile_part;    // [ntiles * 4]: energy, virial, interacting pairs, kinetic energy
    double lp_scale;      // 10^n_prec inside an lp_leapfrog / lp_omf4 integration, 0 otherwise
    double lcheck_dt;     // list MD with rebuild checks every m-th force evaluation: (m - 1) times the
                          // largest drift step of the integration (0: a check before every evaluation)
    // speculative trajectories (the chains are copies of one chain): [0] accepted copy of the last
    // batch (-1: none), [1] trajectories it consumed, [2] copy of its last consumed trajectory; the
    // chain state after the batch, committed by the broadcast
    int* spec_ctl;
    ChainState* spec_st;
    int start_rebuild;    // list MD: an accepted trajectory requests a rebuild at the next trajectory's start
    int lspan;            // list rows are padded with no-op entries past their count up to this many entries
                          // (the sweep shape's unconditional span, at most nbcap)
};

// The updates of lp_leapfrog / lp_omf4 (code-carsten/integrator.hh:51-79, 156-197) round the updated
// momentum or position to the grid 1/s, s = 10^n_prec: v -> round(v s) / s, half away from zero (the
// bit-reversible scheme of Levesque & Verlet 1993); positions are rounded before the wrap.  s == 0:
// the plain update.
__device__ __forceinline__ double lp_round(double val, double s) { return s != 0.0 ? round(val * s) / s : val; }

// Lane exchanges inside a row of 16 lanes through DPP (a VALU move, no LDS round trip): quad_perm
// controls (two bits per lane of a quad), half-row mirror (lane i <- 7 - i) and row mirror (i <- 15 - i)
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;
constexpr int kDppFrom1 = 0xE5, kDppFrom2 = 0xE6;     // lane 0 of a quad <- lane 1 / lane 2 (others keep their own)

// Sum over aligned groups of G lanes (G = 4, 8, 16), in every lane of the group: the xor butterfly
// (steps 1, 2, 4, 8), whose third and fourth steps may take the mirrored lane instead because after the
// previous steps every lane of a quad (half-row) holds the same bits -- additions commute exactly
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v)
{
    auto x = [](T a, auto ctrl) {
        constexpr int c = decltype(ctrl)::value;
        if constexpr (sizeof(T) == 8) return dpp_d<c>(a);
        else return dpp_f<c>(a);
    };
    static_assert(G == 4 || G == 8 || G == 16, "group of 4, 8 or 16 lanes");
    v += x(v, std::integral_constant<int, kDppXor1>{});
    v += x(v, std::integral_constant<int, kDppXor2>{});
    if constexpr (G >= 8) v += x(v, std::integral_constant<int, kDppHalfMirror>{});
    if constexpr (G >= 16) v += x(v, std::integral_constant<int, kDppMirror>{});
    return v;
}

// the periodic wrap x - L floor(x / L) of update_positions (identical_particles.cpp:662-666), with the
// division skipped where it cannot change x: for 0 < x < L the exact quotient is at most 1 - ulp(L) / L,
// more than half a unit below 1, so it rounds below 1, floor gives 0 and the formula returns x itself
__device__ __forceinline__ double wrap_box(double x, double L)
{
#pragma clang fp contract(off)
    return (x > 0.0 && x < L) ? x : x - L * floor(x / L);
}

// LJ_REL64 (list sweep without drift): also the FP64 cell-relative positions of the energy (rel64)
enum LjMode { LJ_FORCE = 1, LJ_KICK = 2, LJ_STOREF = 4, LJ_ENERGY = 8, LJ_DRIFT = 16, LJ_REL64 = 32 };

// one of the two neighbour-list buffers
struct ListBuf {
    char* nb;
    int *nbn, *lperm, *linv, *loffs, *lcs, *ctab;
    float4* lref;         // [Ntot] by slot
    float4* xsl;          // [2 parities][Ntot] by slot
};
__device__ __forceinline__ ListBuf list_buf(const DevView& v, int b)
{
    const size_t N = (size_t)v.Ntot;
    ListBuf L;
    L.nb = v.nb + (size_t)b * v.nb_stride;
    L.nbn = v.nbn + (size_t)b * N;
    L.lperm = v.lperm + (size_t)b * N;
    L.linv = v.linv + (size_t)b * N;
    L.loffs = v.loffs + (size_t)b * ((size_t)v.Cl + 1);
    L.lcs = v.lcs + (size_t)b * N;
    L.ctab = v.ctab + (size_t)b * v.Cl * kSegStride;
    L.lref = v.lref + (size_t)b * N;
    L.xsl = v.xsl + (size_t)b * 2 * N;
    return L;
}
__device__ __forceinline__ int* list_cur(const DevView& v, int chain) { return v.lbuf + chain; }
__device__ __forceinline__ int* list_start(const DevView& v, int chain) { return v.lbuf + v.nchains + chain; }
__device__ __forceinline__ int* list_pord(const DevView& v, int chain) { return v.lbuf + 2 * v.nchains + chain; }
// a thread of chain C that owns slot s marks its tile's current buffer when s begins a tile
__device__ __forceinline__ void tile_cur_mark(const DevView& v, const ChainDev& C, int s, int b)
{
    const int r = s - C.atom_off;
    if (r % kTile == 0) v.tcur[C.tile_first + r / kTile] = b;
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8), and each XCD has
// its own L2: work item xcd_block(b, n) gives XCD x the contiguous range of items
// [x * (n / 8) + min(x, n % 8), ...), so that neighbouring cells or tiles share one L2.  A bijection of
// [0, n) for every n.
__device__ __forceinline__ int xcd_block(int b, int n)
{
    const int x = b & 7, per = n >> 3, rem = n & 7;
    return x * per + min(x, rem) + (b >> 3);
}

// list row of slot s with entries of type E (uint16_t or int, as v.entry16 says)
template <typename E>
__device__ __forceinline__ E* list_row(const DevView& v, const ListBuf& L, int s)
{
    return reinterpret_cast<E*>(L.nb) + (size_t)s * v.nbcap;
}

// 32-bit entries: image code of the neighbour cell offset + 1, two bits per dimension
__device__ __forceinline__ int image_code(int dx, int dy, int dz) { return (dx + 1) | ((dy + 1) << 2) | ((dz + 1) << 4); }
// the offset of one dimension times the cell side, exactly (-side, 0 or +side: f * side - side is exact)
__device__ __forceinline__ float image_shift(int code, int dim, float side)
{
    return __builtin_fmaf((float)((code >> (2 * dim)) & 3), side, -side);
}
constexpr int kImageBits = 6;

// 16-bit entries: cell offset of list segment k in dimension dim (-1, 0, +1)
__host__ __device__ __forceinline__ int seg_offset(int k, int dim) { return (dim == 0 ? k : dim == 1 ? k / 3 : k / 9) % 3 - 1; }
// image shift of segment k in one dimension (the offset times the cell side, exact), or far beyond any
// cut-off for the padding segments
__device__ __forceinline__ float seg_shift(int k, int dim, float side)
{
    return k >= kPadSeg ? kFarShift : (float)seg_offset(k, dim) * side;
}
__device__ __forceinline__ double seg_shift64(int k, int dim, double side)
{
    return k >= kPadSeg ? (double)kFarShift : (double)seg_offset(k, dim) * side;
}
// the first slots of the kSegStride segments of slot s's list cell, copied into an LDS row by the LP
// lanes of the slot (all in one wave); the caller orders these writes before the row's reads
template <int LP>
__device__ __forceinline__ void load_seg_row(const ListBuf& L, const ChainDev& C, int s, int sub, int4* row)
{
    const int4* src = reinterpret_cast<const int4*>(L.ctab + (size_t)(C.lcell_off + L.lcs[s]) * kSegStride);
    for (int w = sub; w < kSegStride / 4; w += LP) row[w] = src[w];
}
// neighbour slot of list entry e, from the slot's segment row
__device__ __forceinline__ int entry_slot(const int* segs, unsigned e) { return segs[e >> kRankBits] + (int)(e & 255u); }

// FP64 position of coordinate x (dimension d) relative to the rebuild cell origin of reference ref,
// taken through the periodic image nearest the rebuild position r0 (an atom moves far less than L / 2
// between rebuilds, so the image is the one of the same box or of a neighbour box)
__device__ __forceinline__ double list_rel(const ChainDev& C, const float4 ref, int d, double x)
{
    const int cell = __float_as_int(ref.w);
    const double o = (double)((cell >> (10 * d)) & 1023) * C.lsize[d];
    const float r0 = d == 0 ? ref.x : (d == 1 ? ref.y : ref.z);
    double r = x - o;
    const double diff = r - (double)r0;
    if (diff > 0.5 * C.L[d]) r -= C.L[d];
    else if (diff < -0.5 * C.L[d]) r += C.L[d];
    return r;
}
__device__ __forceinline__ float ref_coord(const float4 ref, int d) { return d == 0 ? ref.x : (d == 1 ? ref.y : ref.z); }

// current / next binning buffers (the parity of the completed-binning counter selects them)
__device__ __forceinline__ int bin_parity(const DevView& v) { return v.par; }
__device__ __forceinline__ const int* offs_cur(const DevView& v) { return v.offs_buf + bin_parity(v) * (v.Ctot + 1); }
__device__ __forceinline__ int* offs_next(const DevView& v) { return v.offs_buf + (bin_parity(v) ^ 1) * (v.Ctot + 1); }
__device__ __forceinline__ const int* perm_cur(const DevView& v) { return v.perm_buf + bin_parity(v) * v.Ntot; }
__device__ __forceinline__ int* perm_next(const DevView& v) { return v.perm_buf + (bin_parity(v) ^ 1) * v.Ntot; }

// binning: count (cell, slot, new cell at the previous slot) -> scan -> placement, either general
// (scatter + rank by key) or local (collect from the previous neighbour cells)
hipError_t launch_count(const DevView& v, double drift_dt, hipStream_t s);
hipError_t launch_scan(const DevView& v, hipStream_t s);
hipError_t launch_place(const DevView& v, bool local, bool atom_id_order, hipStream_t s);

// per-chain sums of the per-cell (V, W, pairs, K) and per-chunk (K) partials into red[]
enum { SUM_CELLS = 1, SUM_CHUNKS_KI = 2, SUM_CHUNKS_K = 4, SUM_TILES = 8 };
hipError_t launch_reduce(const DevView& v, int what, hipStream_t s);
hipError_t launch_drift(const DevView& v, double dt, hipStream_t s);
hipError_t launch_kick(const DevView& v, double dt, bool fp64, hipStream_t s);
// tiling: i-slots per wave (4, 8, 16) x waves per workgroup (1, 2, 4), encoded islots * 10 + waves
hipError_t launch_lj(const DevView& v, int mode, bool fp64, int tiling, double kick_dt, double drift_dt, hipStream_t s);
hipError_t launch_kinetic(const DevView& v, hipStream_t s);
// slots: write the momenta in the current list buffer's slot order (pl) instead of atom order (p)
hipError_t launch_heatbath(const DevView& v, uint64_t seed, long long traj_or_neg, bool count, bool slots, hipStream_t s);
// speculative batch: chain k draws the momenta of trajectory (device counter + 1 + k) with chain 0's key
hipError_t launch_heatbath_spec(const DevView& v, uint64_t seed, hipStream_t s);
hipError_t launch_metropolis(const DevView& v, const double* uniforms, int ustride, int therm, int ufirst,
                             HmcRecord* rec, int rec_base, int surr, int what, hipStream_t s);
// list_q >= 0: neighbour-list MD at that list parity; saves and restores also switch the list buffers
hipError_t launch_save_restore(const DevView& v, int mode, int list_q, hipStream_t s);
// the same on the slot-ordered state (xl, xl_saved) of list MD, at list parity list_q
hipError_t launch_save_restore_slots(const DevView& v, int mode, int list_q, hipStream_t s);
// speculative batch (the chains are copies of one chain, each ran the next trajectory in turn): the
// Metropolis decisions in copy order up to the first accepted one or trajectory tend (records of the
// consumed trajectories in column 0), save / restore per copy; then the commit of the counter and the
// chain state and the copy of the accepted copy's state into the others (src >= 0: copy chain src
// without a commit)
hipError_t launch_spec_restore(const DevView& v, const double* uniforms, int therm, int ufirst, HmcRecord* rec,
                               int rec_base, int surr, int list_q, hipStream_t s);
hipError_t launch_spec_broadcast(const DevView& v, int src, bool commit, bool rows, hipStream_t s);
hipError_t launch_state_init(const DevView& v, hipStream_t s);
hipError_t launch_spec_final_momenta(const DevView& v, hipStream_t s);
// x, p (and x_saved from xl_saved) <-> xl, pl in the current list buffer's slot order
hipError_t launch_list_import(const DevView& v, hipStream_t s);
hipError_t launch_list_export(const DevView& v, bool saved, hipStream_t s);

// neighbour-list MD: the rebuild (count, [scan,] build; each a no-op unless a chain's request flag of
// parity q is set; cb = the count buffer, alternated by the host at every launch), the FP32
// force sweep (LJ_FORCE | LJ_KICK | LJ_STOREF | LJ_DRIFT; with LJ_DRIFT the parity advances to q ^ 1),
// a standalone drift (q -> q ^ 1) and the FP64 energy / virial / kinetic sums into tile_part
hipError_t launch_list_rebuild(const DevView& v, int q, int cb, hipStream_t s);
hipError_t launch_list_pad_init(const DevView& v, hipStream_t s);
// every row of both list buffers padded with no-op entries from its count up to span
hipError_t launch_list_repad(const DevView& v, int span, hipStream_t s);
// checked: this evaluation's rebuild check ran (its parity's requests are consumed); otherwise a pending
// request is carried to the next parity
hipError_t launch_list_sweep(const DevView& v, int mode, int q, double kick_dt, double drift_dt, bool checked,
                             hipStream_t s);
hipError_t launch_list_drift(const DevView& v, int q, double dt, hipStream_t s);
// rel64: compute the FP64 relative positions first (false when the last sweep wrote them)
hipError_t launch_list_energy(const DevView& v, int q, bool rel64, hipStream_t s);
hipError_t launch_rdf(const DevView& v, int chain, int natoms, int nbins, double size_bin,
                      unsigned long long* counts, hipStream_t s);
hipError_t launch_open_potential(const double* xyz_soa, int n, int nexp, double a, double b,
                                 double* partials, int* nparts, hipStream_t s);
hipError_t launch_open_potential_aos(const double* xyz_aos, int n, int nexp, double a, double b, double* partials,
                                     double* result, hipStream_t s);
hipError_t launch_fill_int(int* p, int value, int n, hipStream_t s);

}  // namespace chmc
