// engine.h - launchers of the device code (host-callable), shared by the
// engine's host code and the kernel translation unit.
#pragma once
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace fnk {

// Slots of the device statistics array.
enum { kStatMaxCell = 0, kStatCellsOver = 1, kStatMaxList = 2, kStatListOver = 3, kStatCount = 4 };

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
    // with chain (a.ch_spring / a.ch_angle layout) evaluating the bonds itself
    static void bead_step(const Args<T, P> &a, bool spr, bool ang, bool next, bool last, bool chain,
                          hipStream_t s);
    // counting sort of the beads by cell: hash (cell, rank in cell, counts),
    // exclusive scan of the counts, place (slots, cell-ordered bead data)
    static void hash(const P *px, const P *py, const P *pz, int n, const Box &b, int *key, int *rank, int *count,
                     hipStream_t s);
    static void place(const int *key, const int *rank, const int *cell_start, const P *px, const P *py, const P *pz,
                      const int *type, const int *tag, const int *loc, int n, int *sorted, const Sorted &so,
                      hipStream_t s);
};

// Neighbour search over the cell-sorted beads into scratch rows of K entries
// per slot, then compaction into the bead-indexed CSR list.
// reach2: the largest squared listing radius of any pair type (a candidate
// farther than that is rejected before its type is read).
void search(const int *sorted, const int *cell_start, int n_cells, const Sorted &so, const double *radius,
            const double *skin, const PairC *pairs, const Box &b, double reach2, int *cnt, int *scratch, int K,
            int *stats, int n, hipStream_t s);
void compact(const int *cnt, const int *nb_start, const int *slot_of, const int *scratch, int K, int n, int *nb_list,
             int cap, hipStream_t s);

void fill_int(int *p, int v, int n, hipStream_t s);
void copy_int(const int *src, int *dst, int n, hipStream_t s);
void add_int(int *p, int v, hipStream_t s);
// Streaming copy of n16 16-byte words (the reference point for traffic counters).
void copy_probe(const void *src, void *dst, size_t n16, hipStream_t s);
int scan_blocks_needed(int n);
// out[0..n] = exclusive prefix sums of in[0..n-1], out[n] = total.  tmp holds
// scan_blocks_needed(n) + 1 ints.
void exclusive_scan(const int *in, int n, int *out, int *tmp, hipStream_t s);

}  // namespace fnk
