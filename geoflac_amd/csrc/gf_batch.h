// gf_batch.h -- many independent models advanced by the same launches.
//
// A parameter sweep holds hundreds of small models of different mesh sizes.  One
// launch per kernel of the step covers all of them: the launch's index range is the
// concatenation of the models' elements (or nodes, thermal tiles, border entries),
// each model's range starting at a multiple of 256.  A workgroup therefore lies in
// one model, which it reads from a table of the model of each 256-index chunk (one
// load, the same for every lane, so the model's entry is read with scalar loads).
// Every model keeps its own parameters, clock (dt, time, reductions) and arrays, so
// the models stay independent; the kernel bodies are the single-model ones.
#pragma once
#include "gf_kernels.h"
#include "gf_phase.h"

struct BatchEntry {
    const GfDevParams *P;
    GfDevClock *ck;
    GfDevArrays A;
    long pe, pn, pt, pb, pm;  // first global index: elements, nodes, thermal tiles, border entries, markers
    long ce, cn, ct, cb, cm;  // counts of the same
};

enum { BK_ELEM = 0, BK_NODE = 1, BK_TILE = 2, BK_BORD = 3, BK_MARK = 4 };
#define BATCH_ALIGN 256  // each model's element, node and border ranges start at a multiple

// model of each 256-index chunk of the element, node and border ranges, and of each tile
struct BatchMap {
    const int *chunk[5];
    int nm, pad;
};

template <int F>
GF_DEV long bprefix(const BatchEntry &b)
{
    return F == BK_ELEM ? b.pe : F == BK_NODE ? b.pn : F == BK_TILE ? b.pt : F == BK_BORD ? b.pb : b.pm;
}
template <int F>
GF_DEV long bcount(const BatchEntry &b)
{
    return F == BK_ELEM ? b.ce : F == BK_NODE ? b.cn : F == BK_TILE ? b.ct : F == BK_BORD ? b.cb : b.cm;
}

// the entry of model m (m uniform across the workgroup), read through the constant
// address space: the table is never written while a kernel runs, so its fields load
// with scalar loads and the buffer descriptors built from them stay uniform (read as
// ordinary global memory, the compiler cannot prove them uniform once the kernel has
// stored anything and wraps every buffer access in a loop over distinct values)
typedef const __attribute__((address_space(4))) BatchEntry ConstBatchEntry;
GF_DEV const BatchEntry &bentry(const BatchEntry *tab, int m)
{
    return *(const BatchEntry *)&((ConstBatchEntry *)tab)[m];
}

// workgroup k is dispatched to XCD k % 8: give each XCD one contiguous run of the
// index range (the first G % 8 XCDs one workgroup more), so that the node rows two
// neighbouring workgroups share stay in one L2 (a bijection for any grid size G)
GF_DEV unsigned xcd_block()
{
    const unsigned G = gridDim.x, k = blockIdx.x, x = k % 8, q = G / 8, r = G % 8;
    return x * q + min(x, r) + k / 8;
}

// calls fn(entry, local index) for the global index of this thread (workgroups of at
// most BATCH_ALIGN threads, dividing it)
template <int F, typename Fn>
GF_DEV void batch_apply(const BatchEntry *__restrict__ tab, const BatchMap &map, Fn fn)
{
    const long g0 = (long)xcd_block() * blockDim.x;
    const int m = __builtin_amdgcn_readfirstlane(map.chunk[F][g0 / BATCH_ALIGN]);
    const BatchEntry &b = bentry(tab, m);
    const long l = g0 + threadIdx.x - bprefix<F>(b);
    if (l >= bcount<F>(b)) return;
    fn(b, l);
}

__global__ void kb_clock(const BatchEntry *__restrict__ tab, BatchMap map, int op)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < map.nm) clock_op(tab[m].ck, op);
}

__global__ void __launch_bounds__(1024) kb_therm_tile(const BatchEntry *__restrict__ tab, BatchMap map, long total,
                                                      int pass)
{
    for (long g = blockIdx.x; g < total; g += gridDim.x) {
        const BatchEntry &b = bentry(tab, __builtin_amdgcn_readfirstlane(map.chunk[BK_TILE][g]));
        if (pass > 0 && b.ck->therm_pass_changed[pass - 1] == 0u) continue;  // this model is proven
        therm_tile_one(b.P, b.ck, b.A, pass, (int)(g - b.pt));
    }
}

// the thermal pipeline over the tiles of all models (k_therm_pipe, one workgroup per
// CU walking a list of tiles); a tile of a model that is already proven only passes
// the prefetch on
template <int XIN, bool DIS>
__global__ void __launch_bounds__(1024) kb_therm_pipe(const BatchEntry *__restrict__ tab, BatchMap map, long total,
                                                      int pass, double stol, int arp)
{
    __shared__ ThermRaw raw;
    __shared__ double sU[TH_NF * TH_NE];
    __shared__ int sFlag[3];
    const long q = (total + 7) / 8, x = blockIdx.x % 8, r = blockIdx.x / 8, RW = gridDim.x / 8;
    const long first = x * q + r, end = min((x + 1) * q, total);
    if (first >= end) return;
    auto entry = [&](long t) -> const BatchEntry & { return bentry(tab, __builtin_amdgcn_readfirstlane(map.chunk[BK_TILE][t])); };
    {
        const BatchEntry &b = entry(first);
        therm_store_raw<XIN, DIS>(therm_prefetch<XIN, DIS>(b.P, b.ck, b.A, pass, (int)(first - b.pt), arp), raw);
    }
    for (long t = first; t < end; t += RW) {
        lds_barrier();  // this tile's inputs are in; the previous tile is done with sU
        const BatchEntry &b = entry(t);
        const long tn = t + RW;
        const BatchEntry &bn = entry(tn < end ? tn : t);
        therm_tile_pipe<XIN, DIS>(b.P, b.ck, b.A, pass, (int)(t - b.pt), bn.P, bn.ck, bn.A,
                                  tn < end ? (int)(tn - bn.pt) : -1, raw, sU, sFlag, stol, arp);
    }
}

__global__ void __launch_bounds__(256) kb_therm_check(const BatchEntry *__restrict__ tab, BatchMap map, int pass)
{
    batch_apply<BK_BORD>(tab, map, [&](const BatchEntry &b, long t) { therm_check_body(b.P, b.ck, b.A, pass, t); });
}

__global__ void __launch_bounds__(256) kb_therm_bc(const BatchEntry *__restrict__ tab, BatchMap map, int passes,
                                                   int proof)
{
    batch_apply<BK_NODE>(tab, map,
                         [&](const BatchEntry &b, long t) { therm_bc_body(b.P, b.ck, b.A, t, passes, proof); });
}

template <int M>
__global__ void __launch_bounds__(256) kb_visc(const BatchEntry *__restrict__ tab, BatchMap map, int rot)
{
    batch_apply<BK_ELEM>(tab, map, [&](const BatchEntry &b, long e) { visc_body<M>(b.P, b.ck, b.A, e, rot); });
}

// rot: the batch's last element move left its stress/strain rotation to this kernel
template <bool SHEARH, int M>
__global__ void __launch_bounds__(256) kb_rheol(const BatchEntry *__restrict__ tab, BatchMap map, int avgsr,
                                                int rot)
{
    batch_apply<BK_ELEM>(tab, map, [&](const BatchEntry &b, long e) {
        rheol_single<SHEARH, typename Prec<M>::SV, typename Prec<M>::AT>(b.P, b.P->ph, b.ck, b.A, e, avgsr, rot);
    });
}

template <int M>
__global__ void __launch_bounds__(256) kb_node(const BatchEntry *__restrict__ tab, BatchMap map)
{
    batch_apply<BK_NODE>(tab, map, [&](const BatchEntry &b, long t) { node_body<M>(b.P, b.ck, b.A, t); });
}

// also records each model's move dt (for a rotation and areas left pending)
template <typename SV>
__global__ void __launch_bounds__(256) kb_move_cord(const BatchEntry *__restrict__ tab, BatchMap map)
{
    batch_apply<BK_NODE>(tab, map, [&](const BatchEntry &b, long t) {
        move_cord_body<SV>(b.P, b.ck, b.A, t);
        if (t == 0) b.ck->dt_move = b.ck->dt;
    });
}

template <int M>
__global__ void __launch_bounds__(256) kb_move_elem(const BatchEntry *__restrict__ tab, BatchMap map)
{
    batch_apply<BK_ELEM>(tab, map, [&](const BatchEntry &b, long e) { move_elem_body<M>(b.P, b.ck, b.A, e); });
}
// the areas, dvol and negative-area STOP of the move (the rotation is left to kb_rheol)
__global__ void __launch_bounds__(256) kb_move_area(const BatchEntry *__restrict__ tab, BatchMap map)
{
    batch_apply<BK_ELEM>(tab, map, [&](const BatchEntry &b, long e) { move_area_body(b.P, b.ck, b.A, e, 0); });
}
// the negative-area STOP of moves whose areas are still pending (coordinates only)
__global__ void __launch_bounds__(256) kb_area_check(const BatchEntry *__restrict__ tab, BatchMap map)
{
    batch_apply<BK_ELEM>(tab, map, [&](const BatchEntry &b, long e) { area_check_body(b.P, b.ck, b.A, e); });
}

template <int M>
__global__ void __launch_bounds__(256) kb_dens(const BatchEntry *__restrict__ tab, BatchMap map, int rot)
{
    batch_apply<BK_ELEM>(tab, map, [&](const BatchEntry &b, long e) { dens_body<M>(b.P, b.ck, b.A, e, rot); });
}

__global__ void __launch_bounds__(256) kb_rmass(const BatchEntry *__restrict__ tab, BatchMap map)
{
    batch_apply<BK_NODE>(tab, map, [&](const BatchEntry &b, long t) { rmass_body(b.P, b.ck, b.A, t); });
}

__global__ void __launch_bounds__(256) kb_amass(const BatchEntry *__restrict__ tab, BatchMap map)
{
    batch_apply<BK_NODE>(tab, map, [&](const BatchEntry &b, long t) { amass_body(b.P, b.ck, b.A, t); });
}

// dt minima: `per` workgroups of 256 threads per model
#define BATCH_DT_BLOCKS 16
template <typename SV>
__global__ void __launch_bounds__(256) kb_dt_reduce(const BatchEntry *__restrict__ tab, BatchMap map)
{
    const int m = blockIdx.x / BATCH_DT_BLOCKS;
    if (m >= map.nm) return;
    const BatchEntry &b = bentry(tab, m);
    dt_reduce_body<SV>(b.P, b.ck, b.A, blockIdx.x % BATCH_DT_BLOCKS, BATCH_DT_BLOCKS);
}

__global__ void kb_dt_scalars(const BatchEntry *__restrict__ tab, BatchMap map)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < map.nm) dt_scalars_body(tab[m].P, tab[m].ck);
}

// change_phase of every model with markers (gf_phase.h)
__global__ void __launch_bounds__(256) kb_phase_rules(const BatchEntry *__restrict__ tab, BatchMap map)
{
    batch_apply<BK_MARK>(tab, map, [&](const BatchEntry &b, long kk) { phase_rules_body(b.P, b.ck, b.A, kk); });
}
// the plan writes the model's active planes into its parameter block
__global__ void kb_phase_plan(const BatchEntry *__restrict__ tab, BatchMap map)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < map.nm && tab[m].A.phstamp) phase_plan_body(const_cast<GfDevParams *>(tab[m].P), tab[m].ck);
}
__global__ void __launch_bounds__(256) kb_phase_replane(const BatchEntry *__restrict__ tab, BatchMap map)
{
    batch_apply<BK_ELEM>(tab, map, [&](const BatchEntry &b, long e) {
        if (b.A.phstamp && b.ck->ph_replan) phase_replane_body(b.P, b.ck, b.A, e);
    });
}
__global__ void __launch_bounds__(256) kb_phase_recount(const BatchEntry *__restrict__ tab, BatchMap map)
{
    batch_apply<BK_ELEM>(tab, map, [&](const BatchEntry &b, long e) {
        if (b.A.phstamp) phase_recount_body(b.P, b.ck, b.A, e);
    });
}
