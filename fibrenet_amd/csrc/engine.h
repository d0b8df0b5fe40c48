// engine.h - launchers of the device code (host-callable), shared by the
// engine's host code and the kernel translation unit.
#pragma once
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace fnk {

// Slots of the device statistics array: cell occupancy (at the last sort),
// the longest listed and candidate lists, search rows that overflowed, the
// candidate builds of the run and their largest total, and the beads not
// searched because their cell held more than kMaxCellBeads beads.
enum {
    kStatMaxCell = 0, kStatCellsOver = 1, kStatMaxList = 2, kStatListOver = 3, kStatMaxCand = 4,
    kStatCandBuilds = 5, kStatCandMax = 6, kStatDense = 7, kStatCount = 8
};

// Beads one grid cell may hold for the searches (the reference keeps 64 per
// cell and drops the rest; beyond this bound a run is refused instead of
// searching a collapsed neighbourhood for minutes).
constexpr int kMaxCellBeads = 4096;

// Cells visited around a bead's cell by the candidate search, per axis: the
// half-width w, or every cell of the axis once (a periodic axis with fewer
// than 2w+1 cells).
struct Stencil {
    int w[3];
    int all[3];
};

// A bead as the rebuild filter reads a partner, one 16-byte record per bead
// written by the displacement check: its position rounded to fp32 and
// FillGrid's cell coordinates packed as cx | cy << sy | cz << sz (-1 outside
// the grid).
struct alignas(16) CellBead {
    float x, y, z;
    int cell;
};

// Bit offsets of cy and cz in CellBead::cell (cx takes sy bits, cy sz - sy).
struct CellPack {
    int sy, sz;
};

// Operands of the rebuild filter.
struct FilterArgs {
    const CellBead *cb;             // rounded positions and cells (disp_check)
    const double *px, *py, *pz;     // the positions themselves
    CellPack pack;
    double delta;                   // bound on the error of a difference of two rounded coordinates
    double cutoff;                  // sqrt of list.cutoff_sq
    double hbs_guard;               // below it, no axis of a pair is wrapped, exactly or rounded
    int n;
    Box grid;                        // FillGrid's settings (cells, cell size)
    Box list;                        // BuildContactList's settings (box, minimum image, cutoff)
    const int *cand_start, *cand;    // candidate CSR
    int *list_start;                 // out: where each bead's list starts (n + 1 entries)
    int cap;                         // entries reserved in cand and list_out
    const PairC *pairs;
    int n_pairs;
    const double *radius, *skin;
    int *list_out, *cnt, *stats;
    int *need;                       // the candidate-build flag, cleared by the filter (null: none)
    const int *net;                  // ensemble: each bead's network, whose Box (boxes[net]) holds its box,
    const Box *boxes;                // grid, cutoff and pair table (at pair_base); null: grid and list above
};

template <typename T, typename P>
struct Launch {
    static void advance_position(const Args<T, P> &a, int n_active, hipStream_t s);
    // bond terms of the step: the first a.n_springs / a.n_angles of the
    // `total` uploaded bonds act, the others leave zero terms
    static void spring_terms(const Args<T, P> &a, int total, hipStream_t s);
    static void angle_terms(const Args<T, P> &a, int total, hipStream_t s);
    // ResolveSprings / ResolveAngles: the terms, then their sum into f of each bead
    static void resolve_springs(const Args<T, P> &a, int total, hipStream_t s);
    static void resolve_angles(const Args<T, P> &a, int total, hipStream_t s);
    static void resolve_contacts(const Args<T, P> &a, hipStream_t s);
    static void advance_velocity(const Args<T, P> &a, int n_active, hipStream_t s);
    static void reset_energy(const Args<T, P> &a, int n_active, hipStream_t s);
    static void drift_first(const Args<T, P> &a, hipStream_t s);
    // the bead kernel of the fast path: after spring_terms / angle_terms, or
    // with chain (a.ch_spring / a.ch_angle layout) evaluating the bonds itself;
    // block: threads per block (128, 256 or 512; the results do not depend on it)
    static void bead_step(const Args<T, P> &a, bool spr, bool ang, bool next, bool last, bool chain, int block,
                          hipStream_t s);
};

// Rebuild.  Kernels taking `gate` do nothing unless *gate is nonzero (gate
// NULL: always run): the candidate build runs only when the displacement
// check asked for it.
//
// Counting sort of the beads by cell: hash writes FillGrid's cell (key, -1
// outside the grid), the cell of the clamped position (ckey), the rank in
// that cell and the counts; an exclusive scan of the counts; place puts the
// beads' data into cell order.
// (net, boxes: the ensemble's network of each bead and Box of each network,
// or NULL for one network described by b)
void hash(const double *px, const double *py, const double *pz, int n, const Box &b, const int *net,
          const Box *boxes, int *key, int *ckey, int *rank, int *count, const int *gate, hipStream_t s);
// (count: the n_count cell counts the offsets were scanned from, zeroed for
// the next sort.)
void place(const int *key, const int *ckey, const int *rank, const int *cell_start, const double *px,
           const double *py, const double *pz, const int *type, const int *tag, const int *loc, int n,
           const Sorted &so, int *count, int n_count, const int *gate, hipStream_t s);
// Partners over the stencil into scratch rows of K entries per bead: the
// candidates within sqrt(r2), or (exact) the pairs BuildContactList lists;
// staged: the neighbourhood of each block read through LDS; flat: the 3x3x3
// stencil's candidates tested by whole waves (k_search_flat; for an
// ensemble, k_search_ens_wave when ens_grids3: every network's grid has at
// least three cells per axis).  Then
// compaction into a bead-indexed CSR sorted per bead (with cnt_out: the final
// list, its lengths and the listed total).
void search(const int *cell_start, int n_cells, const Sorted &so, const PairC *pairs, const double *radius,
            const double *skin, const Box &b, const int *net, const Box *boxes, const Stencil &st, double r2,
            bool ens_grids3, bool exact, int *cnt, int *scratch, int K, int *stats, int n, bool staged, bool flat,
            const int *gate, hipStream_t s);
// Positions to copy into the displacement check's reference when the
// candidates are built (rx null: none).
struct Snapshot {
    const double *px, *py, *pz;
    double *rx, *ry, *rz;
};
// (stats: counts the build and keeps the largest list total of the run;
// snap: the reference positions of the kept candidates.)
void compact(const int *cnt, const int *cand_start, const int *scratch, int K, int n, int *cand,
             int cap, int *cnt_out, int *stats, const Snapshot &snap, const int *gate, hipStream_t s);
// *need = 1 when a bead moved more than sqrt(lim2) since the snapshot; also
// writes every bead's CellBead and FillGrid's cell index (key, -1 outside
// the grid) for the filter and the exports.  With net (an ensemble) each
// bead's network Box stands for both b and grid: CellBead holds the cell
// within the network, key the ensemble's cell index.
void disp_check(const double *px, const double *py, const double *pz, const double *rx, const double *ry,
                const double *rz, int n, const Box &b, const Box &grid, const int *net, const Box *boxes,
                CellPack pack, double lim2, int *need, CellBead *cb, int *key, hipStream_t s);
void snapshot(const double *px, const double *py, const double *pz, double *rx, double *ry, double *rz, int n,
              const int *gate, hipStream_t s);
// The listing rule on the candidates: contact list, counts and cells.
void filter(const FilterArgs &f, hipStream_t s);

// ---- adjoint (adjoint.hip) ----
// Operands of one adjoint step: the forward step's positions and images at
// its force evaluation, its bond statuses after the step (1: the bond acted),
// its contact list, and the fp64 adjoint state.
template <typename T, typename P>
struct AdjArgs {
    int n, n_springs, n_angles;   // bonds at or past n_springs / n_angles did not run
    int springs, angles;          // all bonds (incidence slots 2 * springs, 3 * angles)
    Box drift, force;             // integrator and walls; contacts (single network)
    const int *net;               // batched ensemble: each bead's network, whose Box is boxes[net[i]]
    const Box *boxes;
    const P *px, *py, *pz, *ix, *iy, *iz;
    const T *lam, *hlf, *hlf2, *radius;
    const int *sp_start, *sp_slot, *sp_p1, *sp_p2, *sp_type;
    const unsigned char *sp_on, *an_on;   // per bond: 1 where it acted in the step (status 1), else 0
    const double *sp_l0;
    const T *sp_k;
    const int *an_start, *an_slot, *an_b1, *an_b2, *an_b3, *an_type;
    const T *an_th0, *an_k;
    const int *nb_start, *nb_cnt, *nb_list;
    const PairC *pairs;
    int n_pairs;
    double3 *sp_jv, *an_jv;     // per incidence slot: the bond's J Fb on that bead
    double3 *xb, *wb;           // adjoints of positions and velocity_old
    const double3 *fb;          // adjoint of this step's force (seed and kick term included)
    double3 *fb_next;           // the same for the previous step
    const int *group;           // beads whose x-force is sampled
    const double *seed_prev;    // dL / d(sample) of the previous step per network (null: not sampled)
    double *g_spring, *g_angle;   // adjoints of the force scales of the bonds
    double *g_pair;               // and of the contact pair entries: adjoint_pair_rows rows of n_pairs
                                  // (batched: one row per bead, entries local to its network)
    const int *ch_spring, *ch_angle;   // chain layout (single network): spring (i, i+1), angle at i; or null
    const unsigned *bond_code;         // chain layout: the parameter sets of those bonds (statuses from sp_on /
    const SpringSet *sp_sets;          // an_on); or null: the parameters by bond index
    const AngleSet *an_sets;
};
// One adjoint step: the bond terms (when a.springs / a.angles), then the
// bead update.
template <typename T, typename P>
void adjoint_step(const AdjArgs<T, P> &a, hipStream_t s);
// Fb of the last step: seed[net[i]] (seed[0] without net) at the flagged beads.
template <typename T>
void adjoint_seed(double3 *fb, const double3 *wb, const T *hlf, const int *group, const double *seed,
                  const int *net, int n, hipStream_t s);
// Rows of the contact scale adjoints g_pair (row * n_pairs + entry) that the
// adjoint step accumulates for n beads (chain: bonds in the chain layout).
int adjoint_pair_rows(int n, int np, bool batched, bool chain);
// Whether adjoint_step takes the bonds in its one chain-layout launch (so
// that sp_jv / an_jv are not used).
bool adjoint_one_launch(bool chain, bool batched, int np);
// Batched ensemble: out[pair_base[k] + t] = sum of g[i * cols + t] over the
// beads i in [bead_off[k], bead_off[k + 1]), t < pair_base[k + 1] - pair_base[k].
void sum_network_columns(const double *g, int cols, const int *bead_off, const int *pair_base, int n_net,
                         double *out, hipStream_t s);
// One step's positions (images when im is not null) and bond statuses into
// its record (see k_record_step).
void record_step(const double *const p[3], const double *const im[3], const int *sp_st, int ns, const int *an_st,
                 int na, int n, double *rec_pos, unsigned char *rec_st, hipStream_t s);
// out[t] = sum over the n rows of g[row * np + t]; part holds kSumBlocks * np
// doubles.  Deterministic (fixed partition and order).
constexpr int kSumBlocks = 256;
void sum_columns(const double *g, int n, int np, double *part, double *out, hipStream_t s);
// out[k] = sum of fx[idx[j]] over j in [off[k], off[k + 1]) for k < n_net
// (off on the device, max_len the longest range); part holds
// kSumBlocks * n_net doubles.  Deterministic (fixed partition and order).
template <typename T>
void group_force(const T *fx, const int *idx, const int *off, int n_net, int max_len, double *part, double *out,
                 hipStream_t s);

// The fixed device workspace of fn_group_force_noalloc (noalloc.hip): room
// for kScratchGroup flagged beads over at most kScratchNets networks, with
// the partial sums and results of the reduction; no allocation at any call.
constexpr int kScratchGroup = 1 << 17;
constexpr int kScratchNets = 256;
struct ScratchWorkspace {
    int *idx;       // kScratchGroup flagged bead indices
    int *off;       // kScratchNets + 1 entry offsets per network
    double *part;   // kSumBlocks * kScratchNets partial sums
    double *out;    // kScratchNets results
};
// The workspace of `device`, which must be current: its address is looked up
// at the first call on a device and kept.
hipError_t scratch_workspace(int device, ScratchWorkspace *w);
size_t scratch_bytes();

// The status bits of bond_code[i] from sp_status[ch_spring[i]] and
// an_status[ch_angle[i]] (off without a bond).
void chain_status(const int *ch_spring, const int *ch_angle, const int *sp_status, const int *an_status,
                  unsigned *bond_code, int n, hipStream_t s);
void fill_int(int *p, int v, int n, hipStream_t s);
void fill_gated(int *p, int v, int n, const int *gate, hipStream_t s);
void set_int(int *p, int v, const int *gate, hipStream_t s);
void add_int(int *p, int v, hipStream_t s);
// Streaming copy of n16 16-byte words (the reference point for traffic counters).
void copy_probe(const void *src, void *dst, size_t n16, hipStream_t s);
// dst[0..m) = src[0..m) with m = min(*count, cap) read on the device.
void copy_bounded(const int *src, int *dst, const int *count, int cap, hipStream_t s);
int scan_blocks_needed(int n);
// out[0..n] = exclusive prefix sums of in[0..n-1], out[n] = total.  tmp holds
// scan_blocks_needed(n) ints; in and out are 16-byte aligned.
void exclusive_scan(const int *in, int n, int *out, int *tmp, const int *gate, hipStream_t s);

}  // namespace fnk
