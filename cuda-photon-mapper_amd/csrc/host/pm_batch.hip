// pm_batch.hip -- batched scenes behind the pm_batch_* C ABI.
//
// A frame of a batch is one hipGraph launch of:
//   k_trace_count_batch -> scans -> k_trace_write_batch -> k_tally_final_batch
//   (every scene's photons; scene s owns a block-aligned range of trace
//   blocks, so its records, offsets and tally are those of its own run),
//   kd_build_forest over the 2 x n_scenes photon maps (one level loop for all
//   trees, per-tree sorted lists from one global sort plus a stable pass on
//   the tree id), and one persistent k_render_batch over a device work queue
//   of (scene, tile) items, most expensive scenes first.
// The sizing pass at creation runs the counting part once to learn the record
// counts, so the arenas, the forest tables and the work list are fixed before
// the graph is captured; every replay recomputes the whole frame from scratch.
// No reference counterpart (the reference renders one scene per frame,
// photonMappingKernel.cu:1523-1566).
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <vector>

#include "pm_context.h"

using namespace pm;

struct pm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int ns = 0;
  std::vector<pm_config> cfg;
  DevBuf spheres_d, density_d;
  // emission
  uint32_t tblocks = 0;
  std::vector<BatchTraceScene> tsc;
  DevBuf tsc_d, block_scene_d, cnt_s, cnt_v, off_s, off_v, scan_tmp, small, tally_part, tally9;
  // records: surface arena [0, S_tot) then volume arena [S_tot, S_tot + V_tot)
  uint32_t S_tot = 0, V_tot = 0;
  std::vector<uint32_t> nrec_s, nrec_v, base_s, base_v;
  int bf16 = 0;
  DevBuf x, y, z, payload, tpts;
  DevMap ms, mv;
  // forest: tree t < ns is scene t's surface map, tree ns + t its volume map
  int L = 0;
  uint32_t nnodes = 0;
  std::vector<uint32_t> level_off, tree_nbase, tree_nl, tree_pbase_h, tree_n;
  DevBuf nodes, nrange, point_tree, tree_pbase, tree_root, node_lo, node_hi, node_lc, node_lb, node_child,
      node_pbase, level_nodes, build_tmp;
  // gather mode 7: node boxes and the tree-ordered payload of every tree
  std::vector<KdTreeDesc> tree_desc;
  std::vector<uint32_t> box_level_total;
  uint32_t box_depth = 0, nleaves = 0;
  DevBuf bmin, bmax, tpayload, tree_desc_d, leaf_pref_d, level_pref_d;
  // render
  std::vector<BatchRenderScene> rsc;
  std::vector<uint64_t> img_off;
  uint64_t npix = 0;
  std::vector<uint2> items_h;
  DevBuf rsc_d, items_d, items_tune_d, head, rgb, counters;
  int fp64 = 0, cone = 0, kmax = 1, cus = 1, grid = 1;
  pm_batch_tuning tune;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // last frame
  pm_batch_stats st;
  std::vector<double> tally;
};

namespace {

void release_graph(pm_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

int setup_scenes(pm_batch* b) {
  std::vector<Sphere> sph;
  std::vector<size_t> sph_off(b->ns), dens_off(b->ns, SIZE_MAX);
  std::vector<float> dens;
  for (int i = 0; i < b->ns; ++i) {
    std::vector<Sphere> s = build_spheres(b->cfg[i]);
    sph_off[i] = sph.size();
    sph.insert(sph.end(), s.begin(), s.end());
    std::vector<float> d = build_density(b->cfg[i]);
    if (!d.empty()) {
      dens_off[i] = dens.size();
      dens.insert(dens.end(), d.begin(), d.end());
    }
  }
  HIPCHK(b->spheres_d.ensure(sizeof(Sphere) * std::max<size_t>(sph.size(), 1)));
  if (!sph.empty()) HIPCHK(hipMemcpy(b->spheres_d.p, sph.data(), sizeof(Sphere) * sph.size(), hipMemcpyHostToDevice));
  if (!dens.empty()) {
    HIPCHK(b->density_d.ensure(sizeof(float) * dens.size()));
    HIPCHK(hipMemcpy(b->density_d.p, dens.data(), sizeof(float) * dens.size(), hipMemcpyHostToDevice));
  }
  const uint32_t TT = trace_threads();
  b->tsc.resize(b->ns);
  b->tblocks = 0;
  std::vector<uint32_t> block_scene;
  for (int i = 0; i < b->ns; ++i) {
    const pm_config& c = b->cfg[i];
    BatchTraceScene& t = b->tsc[i];
    memset(&t, 0, sizeof(t));
    t.sc = make_scene(c, b->spheres_d.as<Sphere>() + sph_off[i],
                      dens_off[i] == SIZE_MAX ? nullptr : b->density_d.as<float>() + dens_off[i]);
    t.tp = trace_params(c);
    t.n_emit = c.n_emit;
    t.block0 = b->tblocks;
    t.nblocks = ceil_div(c.n_emit, TT);
    b->tblocks += t.nblocks;
    block_scene.insert(block_scene.end(), t.nblocks, (uint32_t)i);
  }
  HIPCHK(b->tsc_d.ensure(sizeof(BatchTraceScene) * b->ns));
  HIPCHK(hipMemcpy(b->tsc_d.p, b->tsc.data(), sizeof(BatchTraceScene) * b->ns, hipMemcpyHostToDevice));
  HIPCHK(b->block_scene_d.ensure(4ull * b->tblocks));
  HIPCHK(hipMemcpy(b->block_scene_d.p, block_scene.data(), 4ull * b->tblocks, hipMemcpyHostToDevice));
  return PM_OK;
}

// Counting pass outside the graph: per-scene record counts and bases; arenas.
int size_records(pm_batch* b) {
  hipStream_t st = b->stream;
  const size_t np = (size_t)b->tblocks * trace_threads();
  HIPCHK(b->cnt_s.ensure(4 * np));
  HIPCHK(b->cnt_v.ensure(4 * np));
  HIPCHK(b->off_s.ensure(4 * np));
  HIPCHK(b->off_v.ensure(4 * np));
  HIPCHK(b->scan_tmp.ensure(4 * scan_temp_words(np)));
  HIPCHK(b->small.ensure(64));
  uint32_t* small = b->small.as<uint32_t>();
  HIPCHK(launch_trace_count_batch(b->tsc_d.as<BatchTraceScene>(), b->block_scene_d.as<uint32_t>(), b->tblocks,
                                  b->cnt_s.as<uint32_t>(), b->cnt_v.as<uint32_t>(), st));
  HIPCHK(exclusive_scan(b->cnt_s.as<uint32_t>(), b->off_s.as<uint32_t>(), np, b->scan_tmp.as<uint32_t>(), small, st));
  HIPCHK(exclusive_scan(b->cnt_v.as<uint32_t>(), b->off_v.as<uint32_t>(), np, b->scan_tmp.as<uint32_t>(), small + 1,
                        st));
  std::vector<uint32_t> first(b->ns);
  for (int i = 0; i < b->ns; ++i) first[i] = b->tsc[i].block0 * trace_threads();
  DevBuf idx, got;
  HIPCHK(idx.ensure(4ull * b->ns));
  HIPCHK(got.ensure(8ull * b->ns));
  HIPCHK(hipMemcpyAsync(idx.p, first.data(), 4ull * b->ns, hipMemcpyHostToDevice, st));
  HIPCHK(launch_gather_u32(b->off_s.as<uint32_t>(), idx.as<uint32_t>(), b->ns, got.as<uint32_t>(), st));
  HIPCHK(launch_gather_u32(b->off_v.as<uint32_t>(), idx.as<uint32_t>(), b->ns, got.as<uint32_t>() + b->ns, st));
  std::vector<uint32_t> h(2 * b->ns);
  uint32_t tot[2];
  HIPCHK(hipMemcpyAsync(h.data(), got.p, 8ull * b->ns, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(tot, small, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  idx.release();
  got.release();
  if ((uint64_t)tot[0] + tot[1] >= (1ull << 31)) return PM_ERR_CAPACITY;
  b->S_tot = tot[0];
  b->V_tot = tot[1];
  b->base_s.assign(h.begin(), h.begin() + b->ns);
  b->base_v.assign(h.begin() + b->ns, h.end());
  b->nrec_s.resize(b->ns);
  b->nrec_v.resize(b->ns);
  for (int i = 0; i < b->ns; ++i) {
    b->nrec_s[i] = (i + 1 < b->ns ? b->base_s[i + 1] : b->S_tot) - b->base_s[i];
    b->nrec_v[i] = (i + 1 < b->ns ? b->base_v[i + 1] : b->V_tot) - b->base_v[i];
  }
  const size_t N = std::max<size_t>((size_t)b->S_tot + b->V_tot, 1);
  const size_t stride = b->bf16 ? 16 : 32;
  HIPCHK(b->x.ensure(4 * N));
  HIPCHK(b->y.ensure(4 * N));
  HIPCHK(b->z.ensure(4 * N));
  HIPCHK(b->payload.ensure(stride * N));
  HIPCHK(b->tpts.ensure(16 * (N + kTptsPad)));
  for (int v = 0; v < 2; ++v) {
    DevMap& m = v ? b->mv : b->ms;
    memset(&m, 0, sizeof(m));
    const size_t base = v ? b->S_tot : 0;
    m.n = m.cap = v ? b->V_tot : b->S_tot;
    m.bf16 = b->bf16;
    m.x = b->x.as<float>() + base;
    m.y = b->y.as<float>() + base;
    m.z = b->z.as<float>() + base;
    m.payload = (char*)b->payload.p + stride * base;
    m.tpts = b->tpts.as<float4>() + base;
  }
  HIPCHK(b->tally_part.ensure(sizeof(double) * 9 * std::max<uint32_t>(b->tblocks, 1)));
  HIPCHK(b->tally9.ensure(sizeof(double) * 9 * b->ns));
  return PM_OK;
}

GatherMap tree_gather_map(const pm_batch* b, int scene, int vol, uint32_t tree) {
  const pm_config& c = b->cfg[scene];
  GatherMap g;
  memset(&g, 0, sizeof(g));
  const size_t stride = b->bf16 ? 16 : 32;
  g.tpts = b->tpts.as<float4>() + b->tree_pbase_h[tree];
  g.nodes = b->nodes.as<KdNode>() + b->tree_nbase[tree];
  g.nrange = b->nrange.as<uint2>() + b->tree_nbase[tree];
  g.payload = (const char*)b->payload.p + stride * b->tree_pbase_h[tree];
  g.bmin = b->bmin.as<float4>() + b->tree_desc[tree].bbase;
  g.bmax = b->bmax.as<float4>() + b->tree_desc[tree].bbase;
  // the tree-ordered payload exists once gather mode 7 has been chosen
  g.tpayload = b->tpayload.p ? (const char*)b->tpayload.p + stride * b->tree_pbase_h[tree] : nullptr;
  g.n = b->tree_n[tree];
  g.nl = b->tree_nl[tree];
  g.bf16 = b->bf16;
  g.k = vol ? c.k_volume : c.k_surface;
  const float rc = vol ? c.rcap_volume : c.rcap_surface;
  g.rcap2 = rc * rc;
  for (int j = 0; j < 3; ++j) g.scale[j] = c.light_power[j] / (float)c.n_emit;
  return g;
}

int setup_scene_maps(pm_batch* b);

// Forest tables for leaf size L, the per-scene gather maps, the build scratch.
int setup_forest(pm_batch* b, int L) {
  const uint32_t nt = 2u * (uint32_t)b->ns;
  b->tree_n.resize(nt);
  b->tree_pbase_h.resize(nt);
  b->tree_nl.resize(nt);
  b->tree_nbase.resize(nt);
  std::vector<std::vector<uint32_t>> lo(nt), hi(nt), lc(nt), lbl(nt);
  uint32_t nn = 0, depth = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    const int s = (int)(t % (uint32_t)b->ns);
    const bool vol = t >= (uint32_t)b->ns;
    b->tree_n[t] = vol ? b->nrec_v[s] : b->nrec_s[s];
    b->tree_pbase_h[t] = vol ? b->S_tot + b->base_v[s] : b->base_s[s];
    b->tree_nl[t] = kd_num_leaves(b->tree_n[t], (uint32_t)L);
    b->tree_nbase[t] = nn;
    shape_tables(b->tree_n[t], (uint32_t)L, lo[t], hi[t], lc[t], lbl[t]);
    const uint32_t ni = (uint32_t)lo[t].size();
    nn += ni;
    uint32_t d = 0;
    for (uint64_t f = 0; f < ni; f = 2 * f + 1) ++d;
    depth = std::max(depth, d);
  }
  b->nnodes = nn;
  std::vector<uint32_t> nlo(nn), nhi(nn), nlc(nn), nlb(nn), nchild(2ull * nn), npb(nn), lvl, root(nt), pbase(nt);
  b->level_off.assign(1, 0u);
  for (uint32_t l = 0; l < depth; ++l) {
    uint32_t run = 0;
    const uint64_t f = (1ull << l) - 1ull, e = (2ull << l) - 1ull;
    for (uint32_t t = 0; t < nt; ++t) {  // trees in point order
      const uint32_t ni = (uint32_t)lo[t].size(), nb = b->tree_nbase[t], pb = b->tree_pbase_h[t];
      for (uint64_t h = f; h < std::min<uint64_t>(e, ni); ++h) {
        const uint32_t id = nb + (uint32_t)h;
        nlo[id] = pb + lo[t][h];
        nhi[id] = pb + hi[t][h];
        nlc[id] = lc[t][h];
        nlb[id] = run;
        run += lc[t][h];
        const uint64_t cl = 2 * h + 1, cr = 2 * h + 2;
        nchild[2ull * id] = cl < ni ? nb + (uint32_t)cl : KD_LEAF;
        nchild[2ull * id + 1] = cr < ni ? nb + (uint32_t)cr : KD_LEAF;
        npb[id] = pb;
        lvl.push_back(id);
      }
    }
    b->level_off.push_back((uint32_t)lvl.size());
  }
  for (uint32_t t = 0; t < nt; ++t) {
    root[t] = lo[t].empty() ? KD_LEAF : b->tree_nbase[t];
    pbase[t] = b->tree_pbase_h[t];
  }
  const size_t N = (size_t)b->S_tot + b->V_tot;
  const size_t n1 = std::max<size_t>(nn, 1);
  HIPCHK(b->nodes.ensure(sizeof(KdNode) * n1));
  HIPCHK(b->nrange.ensure(sizeof(uint2) * n1));
  DevBuf* tabs[6] = {&b->node_lo, &b->node_hi, &b->node_lc, &b->node_lb, &b->node_pbase, &b->level_nodes};
  const std::vector<uint32_t>* src[6] = {&nlo, &nhi, &nlc, &nlb, &npb, &lvl};
  for (int i = 0; i < 6; ++i) {
    HIPCHK(tabs[i]->ensure(4 * n1));
    if (!src[i]->empty())
      HIPCHK(hipMemcpy(tabs[i]->p, src[i]->data(), 4 * src[i]->size(), hipMemcpyHostToDevice));
  }
  HIPCHK(b->node_child.ensure(8 * n1));
  if (nn) HIPCHK(hipMemcpy(b->node_child.p, nchild.data(), 8ull * nn, hipMemcpyHostToDevice));
  HIPCHK(b->tree_root.ensure(4ull * nt));
  HIPCHK(b->tree_pbase.ensure(4ull * nt));
  HIPCHK(hipMemcpy(b->tree_root.p, root.data(), 4ull * nt, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->tree_pbase.p, pbase.data(), 4ull * nt, hipMemcpyHostToDevice));
  HIPCHK(b->point_tree.ensure(4 * std::max<size_t>(N, 1)));
  for (uint32_t t = 0; t < nt; ++t)
    HIPCHK(launch_fill_u32(b->point_tree.as<uint32_t>() + b->tree_pbase_h[t], t, b->tree_n[t], b->stream));
  HIPCHK(b->build_tmp.ensure(4 * kdforest_temp_words((uint32_t)N, nn)));
  // node boxes (gather mode 7): 2 nl - 1 per tree, leaves and internal levels
  b->tree_desc.resize(nt);
  std::vector<uint32_t> lpref(nt + 1, 0u);
  uint32_t nbox = 0, bdepth = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    KdTreeDesc& D = b->tree_desc[t];
    D.n = b->tree_n[t];
    D.nl = b->tree_nl[t];
    D.nbase = b->tree_nbase[t];
    D.pbase = b->tree_pbase_h[t];
    D.bbase = nbox;
    D.pad = 0;
    nbox += D.nl ? 2u * D.nl - 1u : 0u;
    lpref[t + 1] = lpref[t] + D.nl;
    const uint32_t ni = D.nl ? D.nl - 1u : 0u;
    uint32_t d = 0;
    while (d < 32u && (1ull << d) - 1ull < ni) ++d;  // levels 0 .. d - 1 hold internal nodes
    bdepth = std::max(bdepth, d);
  }
  std::vector<uint32_t> lvpref((size_t)bdepth * (nt + 1), 0u);
  b->box_level_total.assign(bdepth, 0u);
  for (uint32_t d = 0; d < bdepth; ++d) {
    uint32_t* row = lvpref.data() + (size_t)d * (nt + 1);
    for (uint32_t t = 0; t < nt; ++t) {
      const uint32_t ni = b->tree_desc[t].nl ? b->tree_desc[t].nl - 1u : 0u;
      const uint64_t first = (1ull << d) - 1ull, last = std::min<uint64_t>((2ull << d) - 2ull, (uint64_t)ni - 1ull);
      const uint32_t c = ni > first ? (uint32_t)(last - first + 1ull) : 0u;
      row[t + 1] = row[t] + c;
    }
    b->box_level_total[d] = row[nt];
  }
  b->box_depth = bdepth;
  b->nleaves = lpref[nt];
  HIPCHK(b->bmin.ensure(16ull * std::max<uint32_t>(nbox, 1)));
  HIPCHK(b->bmax.ensure(16ull * std::max<uint32_t>(nbox, 1)));
  HIPCHK(b->tree_desc_d.ensure(sizeof(KdTreeDesc) * nt));
  HIPCHK(hipMemcpy(b->tree_desc_d.p, b->tree_desc.data(), sizeof(KdTreeDesc) * nt, hipMemcpyHostToDevice));
  HIPCHK(b->leaf_pref_d.ensure(4ull * (nt + 1)));
  HIPCHK(hipMemcpy(b->leaf_pref_d.p, lpref.data(), 4ull * (nt + 1), hipMemcpyHostToDevice));
  HIPCHK(b->level_pref_d.ensure(4ull * std::max<size_t>(lvpref.size(), 1)));
  if (!lvpref.empty())
    HIPCHK(hipMemcpy(b->level_pref_d.p, lvpref.data(), 4ull * lvpref.size(), hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->L = L;
  return setup_scene_maps(b);
}

// Per-scene render descriptors (the gather maps point into the forest and,
// for gather mode 7, into the tree-ordered payload).
int setup_scene_maps(pm_batch* b) {
  b->rsc.resize(b->ns);
  for (int s = 0; s < b->ns; ++s) {
    const pm_config& c = b->cfg[s];
    BatchRenderScene& r = b->rsc[s];
    memset(&r, 0, sizeof(r));
    r.sc = b->tsc[s].sc;
    r.rp = render_params(c);
    r.gs = tree_gather_map(b, s, 0, (uint32_t)s);
    r.gv = tree_gather_map(b, s, 1, (uint32_t)(b->ns + s));
    r.exposure = c.exposure;
    r.w = c.width;
    r.h = c.height;
    r.img_off = b->img_off[s];
  }
  HIPCHK(b->rsc_d.ensure(sizeof(BatchRenderScene) * b->ns));
  HIPCHK(hipMemcpy(b->rsc_d.p, b->rsc.data(), sizeof(BatchRenderScene) * b->ns, hipMemcpyHostToDevice));
  return PM_OK;
}

// Image arena and the work list: the 8x8 tiles of every scene, scenes in
// decreasing order of expected cost per tile (march length x log of the map
// sizes x k), so the persistent grid ends on the cheapest tiles.
int setup_render(pm_batch* b) {
  b->img_off.resize(b->ns);
  b->npix = 0;
  std::vector<double> cost(b->ns);
  std::vector<int> order(b->ns);
  for (int s = 0; s < b->ns; ++s) {
    const pm_config& c = b->cfg[s];
    b->img_off[s] = b->npix;
    b->npix += (uint64_t)c.width * (uint64_t)c.height;
    const bool media = c.media_on && c.medium != PM_MEDIUM_NONE;
    cost[s] = c.spp * ((double)c.k_surface * std::log2(b->nrec_s[s] + 2.0) +
                       (media ? (double)c.k_volume * std::log2(b->nrec_v[s] + 2.0) * 1.5 / c.march_step : 0.0));
    order[s] = s;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
  b->items_h.clear();
  for (int s : order) {
    const pm_config& c = b->cfg[s];
    const uint32_t tiles = ceil_div(c.width, 8) * ceil_div(c.height, 8);
    for (uint32_t t = 0; t < tiles; ++t) b->items_h.push_back(make_uint2((uint32_t)s, t));
  }
  HIPCHK(b->items_d.ensure(sizeof(uint2) * std::max<size_t>(b->items_h.size(), 1)));
  HIPCHK(hipMemcpy(b->items_d.p, b->items_h.data(), sizeof(uint2) * b->items_h.size(), hipMemcpyHostToDevice));
  HIPCHK(b->rgb.ensure(12 * std::max<uint64_t>(b->npix, 1)));
  HIPCHK(b->head.ensure(16));
  HIPCHK(b->counters.ensure(16));
  return PM_OK;
}

int set_grid(pm_batch* b) {
  const int occ = render_batch_occupancy(b->fp64, b->cone, b->tune.block_threads, b->tune.heap_slots,
                                         b->tune.heap_key_bytes, b->tune.gather_mode == 7, b->tune.top_levels);
  if (occ < 1) return PM_ERR_INVALID;
  b->grid = b->cus * occ;
  return PM_OK;
}

// Enqueues one frame over `nitems` work items on b->stream.  ev (optional,
// outside graph capture only) receives events after the emission and after
// the build, so a caller can time the phases.
hipError_t enqueue_frame(pm_batch* b, const uint2* items, uint32_t nitems, hipEvent_t ev_emitted,
                         hipEvent_t ev_built) {
  hipStream_t st = b->stream;
  const size_t np = (size_t)b->tblocks * trace_threads();
  uint32_t* small = b->small.as<uint32_t>();
  KdForest F;
  memset(&F, 0, sizeof(F));
  F.n = b->S_tot + b->V_tot;
  F.x = b->x.as<float>(); F.y = b->y.as<float>(); F.z = b->z.as<float>();
  F.tpts = b->tpts.as<float4>();
  F.nodes = b->nodes.as<KdNode>();
  F.nrange = b->nrange.as<uint2>();
  F.ntrees = 2u * (uint32_t)b->ns;
  F.nnodes = b->nnodes;
  F.point_tree = b->point_tree.as<uint32_t>();
  F.tree_pbase = b->tree_pbase.as<uint32_t>();
  F.tree_root = b->tree_root.as<uint32_t>();
  F.node_lo = b->node_lo.as<uint32_t>();
  F.node_hi = b->node_hi.as<uint32_t>();
  F.node_lc = b->node_lc.as<uint32_t>();
  F.node_lb = b->node_lb.as<uint32_t>();
  F.node_child = b->node_child.as<uint32_t>();
  F.node_pbase = b->node_pbase.as<uint32_t>();
  F.level_nodes = b->level_nodes.as<uint32_t>();
  BatchRenderLaunch R;
  memset(&R, 0, sizeof(R));
  R.scenes = b->rsc_d.as<BatchRenderScene>();
  R.items = items;
  R.nitems = nitems;
  R.head = b->head.as<uint32_t>();
  R.rgb = b->rgb.as<float>();
  R.counters = b->counters.as<unsigned long long>();
  R.heap_slots = b->tune.heap_slots;
  R.key_bytes = b->tune.heap_key_bytes;
  R.counting = b->tune.gather_mode == 7;
  R.top_levels = b->tune.top_levels;
  hipError_t e;
  if ((e = launch_trace_count_batch(b->tsc_d.as<BatchTraceScene>(), b->block_scene_d.as<uint32_t>(), b->tblocks,
                                    b->cnt_s.as<uint32_t>(), b->cnt_v.as<uint32_t>(), st)) != hipSuccess)
    return e;
  if ((e = exclusive_scan(b->cnt_s.as<uint32_t>(), b->off_s.as<uint32_t>(), np, b->scan_tmp.as<uint32_t>(), small,
                          st)) != hipSuccess)
    return e;
  if ((e = exclusive_scan(b->cnt_v.as<uint32_t>(), b->off_v.as<uint32_t>(), np, b->scan_tmp.as<uint32_t>(),
                          small + 1, st)) != hipSuccess)
    return e;
  if ((e = launch_trace_write_batch(b->tsc_d.as<BatchTraceScene>(), b->block_scene_d.as<uint32_t>(), b->tblocks,
                                    b->off_s.as<uint32_t>(), b->off_v.as<uint32_t>(), b->ms, b->mv,
                                    b->tally_part.as<double>(), st)) != hipSuccess)
    return e;
  if ((e = launch_tally_final_batch(b->tally_part.as<double>(), b->tsc_d.as<BatchTraceScene>(), (uint32_t)b->ns,
                                    b->tally9.as<double>(), st)) != hipSuccess)
    return e;
  if (ev_emitted && (e = hipEventRecord(ev_emitted, st)) != hipSuccess) return e;
  if ((e = kd_build_forest(F, b->level_off.data(), (uint32_t)b->level_off.size() - 1u, b->build_tmp.as<uint32_t>(),
                           st)) != hipSuccess)
    return e;
  {  // node boxes (every gather prunes with them); the tree-ordered payload for mode 7
    KdForestBoxes fb;
    fb.tpts = F.tpts;
    fb.nodes = F.nodes;
    fb.nrange = F.nrange;
    fb.trees = b->tree_desc_d.as<KdTreeDesc>();
    fb.ntrees = F.ntrees;
    fb.leaf_pref = b->leaf_pref_d.as<uint32_t>();
    fb.level_pref = b->level_pref_d.as<uint32_t>();
    fb.depth = b->box_depth;
    fb.nleaves = b->nleaves;
    fb.bmin = b->bmin.as<float4>();
    fb.bmax = b->bmax.as<float4>();
    if ((e = kd_forest_boxes(fb, b->box_level_total.data(), st)) != hipSuccess) return e;
    if (R.counting &&
        (e = kd_forest_payload(F.tpts, F.point_tree, F.tree_pbase, b->payload.p, b->tpayload.p, F.n, b->bf16, st)) !=
            hipSuccess)
      return e;
  }
  if (ev_built && (e = hipEventRecord(ev_built, st)) != hipSuccess) return e;
  if ((e = launch_fill_u32(b->head.as<uint32_t>(), 0u, 4, st)) != hipSuccess) return e;
  if ((e = launch_fill_u32(b->counters.as<uint32_t>(), 0u, 4, st)) != hipSuccess) return e;
  return launch_render_batch(R, b->fp64, b->cone, b->tune.block_threads, b->grid, st);
}

// Captures one frame over every work item into b->graph / b->exec.
int capture(pm_batch* b) {
  release_graph(b);
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = enqueue_frame(b, b->items_d.as<uint2>(), (uint32_t)b->items_h.size(), nullptr, nullptr);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    set_last_hip(e != hipSuccess ? e : e2, "batch graph capture");
    return PM_ERR_HIP;
  }
  b->graph = g;
  HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  size_t n = 0;
  HIPCHK(hipGraphGetNodes(b->graph, nullptr, &n));
  b->st.graph_nodes = (uint32_t)n;
  return PM_OK;
}

int run_frame(pm_batch* b, double* ms) {
  HIPCHK(hipEventRecord(b->e0, b->stream));
  HIPCHK(hipGraphLaunch(b->exec, b->stream));
  HIPCHK(hipEventRecord(b->e1, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  float t = 0.0f;
  HIPCHK(hipEventElapsedTime(&t, b->e0, b->e1));
  *ms = t;
  return PM_OK;
}

// capture_graph = false leaves the graph to the caller (the autotuner runs
// the frame directly on the stream).
int apply_tuning(pm_batch* b, const pm_batch_tuning& t, bool capture_graph) {
  if (t.leaf_size < 1 || t.leaf_size > 64) return PM_ERR_INVALID;
  if (t.block_threads != 64 && t.block_threads != 128 && t.block_threads != 192 && t.block_threads != 256)
    return PM_ERR_INVALID;
  if ((t.heap_slots != 32 && t.heap_slots != 64) || t.heap_slots < b->kmax) return PM_ERR_INVALID;
  if (t.heap_key_bytes != 8 && t.heap_key_bytes != 4 && t.heap_key_bytes != 2) return PM_ERR_INVALID;
  if (t.gather_mode != 0 && t.gather_mode != 7) return PM_ERR_INVALID;
  if (t.gather_mode == 7 && (t.heap_key_bytes != 2 || t.heap_slots != 64)) return PM_ERR_INVALID;
  if (t.top_levels < 0 || t.top_levels > kMaxTopLevels || (t.top_levels > 0 && t.gather_mode != 7))
    return PM_ERR_INVALID;
  int r;
  // the grid is checked first, then the tuning is committed; a later failure
  // puts the previous tuning (and its grid) back
  const pm_batch_tuning prev = b->tune;
  const int prev_grid = b->grid;
  b->tune = t;
  if ((r = set_grid(b)) != PM_OK) {
    b->tune = prev;
    b->grid = prev_grid;
    return r;
  }
  auto restore = [&](int err) {
    b->tune = prev;
    b->grid = prev_grid;
    return err;
  };
  if (t.gather_mode == 7 && !b->tpayload.p) {
    // the tree-ordered payload of gather mode 7, allocated on first use
    const size_t N = std::max<size_t>((size_t)b->S_tot + b->V_tot, 1);
    if (b->tpayload.ensure((b->bf16 ? 16 : 32) * N) != hipSuccess) return restore(PM_ERR_NOMEM);
    if (t.leaf_size == b->L && (r = setup_scene_maps(b)) != PM_OK) return restore(r);
  }
  if (t.leaf_size != b->L && (r = setup_forest(b, t.leaf_size)) != PM_OK) return restore(r);
  if (!capture_graph) {
    release_graph(b);
    return PM_OK;
  }
  return capture(b);
}

void destroy(pm_batch* b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  release_graph(b);
  for (DevBuf* d : {&b->spheres_d, &b->density_d, &b->tsc_d, &b->block_scene_d, &b->cnt_s, &b->cnt_v, &b->off_s,
                    &b->off_v, &b->scan_tmp, &b->small, &b->tally_part, &b->tally9, &b->x, &b->y, &b->z,
                    &b->payload, &b->tpts, &b->nodes, &b->nrange, &b->point_tree, &b->tree_pbase, &b->tree_root,
                    &b->node_lo, &b->node_hi, &b->node_lc, &b->node_lb, &b->node_child, &b->node_pbase,
                    &b->level_nodes, &b->build_tmp, &b->rsc_d, &b->items_d, &b->items_tune_d, &b->head, &b->rgb,
                    &b->counters, &b->bmin, &b->bmax, &b->tpayload, &b->tree_desc_d, &b->leaf_pref_d,
                    &b->level_pref_d})
    d->release();
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

bool scene_ok(const pm_batch* b, int scene) { return b && scene >= 0 && scene < b->ns; }

DevMap scene_map(const pm_batch* b, int scene, int map) {
  const uint32_t t = (uint32_t)scene + (map == PM_MAP_VOLUME ? (uint32_t)b->ns : 0u);
  const DevMap& a = map == PM_MAP_VOLUME ? b->mv : b->ms;
  const size_t off = map == PM_MAP_VOLUME ? b->base_v[scene] : b->base_s[scene];
  const size_t stride = b->bf16 ? 16 : 32;
  DevMap m = a;
  m.n = m.cap = b->tree_n[t];
  m.x = a.x + off;
  m.y = a.y + off;
  m.z = a.z + off;
  m.payload = (char*)a.payload + stride * off;
  m.tpts = b->tpts.as<float4>() + b->tree_pbase_h[t];
  m.nodes = b->nodes.as<KdNode>() + b->tree_nbase[t];
  m.nrange = b->nrange.as<uint2>() + b->tree_nbase[t];
  m.L = (uint32_t)b->L;
  m.nl = b->tree_nl[t];
  return m;
}

}  // namespace

extern "C" {

int pm_batch_scene_config(uint32_t seed, int index, pm_config* out) {
  if (!out) return PM_ERR_INVALID;
  return batch_scene_config(seed, index, out) == 0 ? PM_OK : PM_ERR_INVALID;
}

int pm_batch_create(const pm_config* cfgs, int n, pm_batch** out) {
  if (!cfgs || !out || n < 1 || n > 128) return PM_ERR_INVALID;
  uint64_t max_records = 0;
  for (int i = 0; i < n; ++i) {
    const pm_config& c = cfgs[i];
    if (validate_config(c) != 0 || c.n_emit == 0 || c.cap_surface || c.cap_volume) return PM_ERR_INVALID;
    if (c.storage != cfgs[0].storage || c.estimator != cfgs[0].estimator ||
        c.fp64_transmittance != cfgs[0].fp64_transmittance)
      return PM_ERR_INVALID;
    max_records += (uint64_t)c.n_emit * (uint64_t)c.max_bounces;
  }
  // record offsets of the batch arena are 32-bit, and the arena holds fewer
  // than 2^31 records: refuse any batch whose photons could store more
  if (max_records >= (1ull << 31)) return PM_ERR_CAPACITY;
  pm_batch* b = new pm_batch();
  b->ns = n;
  b->cfg.assign(cfgs, cfgs + n);
  b->bf16 = cfgs[0].storage == PM_STORE_BF16;
  b->fp64 = cfgs[0].fp64_transmittance != 0;
  b->cone = cfgs[0].estimator == PM_EST_CONE;
  for (int i = 0; i < n; ++i) b->kmax = std::max(b->kmax, std::max(cfgs[i].k_surface, cfgs[i].k_volume));
  memset(&b->st, 0, sizeof(b->st));
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->e0);
  if (e == hipSuccess) e = hipEventCreate(&b->e1);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device);
  if (e != hipSuccess) {
    set_last_hip(e, "pm_batch_create");
    destroy(b);
    return PM_ERR_HIP;
  }
  const auto t0 = std::chrono::steady_clock::now();
  pm_batch_tuning t;
  t.leaf_size = 32;
  t.block_threads = 64;
  t.heap_slots = b->kmax <= 32 ? 32 : 64;
  t.heap_key_bytes = 2;
  t.gather_mode = 0;
  t.top_levels = 0;
  int r = setup_scenes(b);
  if (r == PM_OK) r = size_records(b);
  if (r == PM_OK) r = setup_render(b);
  if (r == PM_OK) r = apply_tuning(b, t, true);
  if (r != PM_OK) {
    destroy(b);
    return r;
  }
  b->st.ms_setup = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  b->st.n_scenes = (uint32_t)n;
  b->st.work_items = (uint32_t)b->items_h.size();
  for (int i = 0; i < n; ++i) b->st.photons += cfgs[i].n_emit;
  b->st.surface_records = b->S_tot;
  b->st.volume_records = b->V_tot;
  b->st.pixels = b->npix;
  b->tally.assign(9ull * n, 0.0);
  *out = b;
  return PM_OK;
}

int pm_batch_destroy(pm_batch* b) {
  if (!b) return PM_ERR_INVALID;
  destroy(b);
  return PM_OK;
}

int pm_batch_set_tuning(pm_batch* b, const pm_batch_tuning* t) {
  if (!b || !t) return PM_ERR_INVALID;
  return apply_tuning(b, *t, true);
}

int pm_batch_autotune(pm_batch* b, int sample_stride, pm_batch_tuning* best_out, double* best_ms_out) {
  if (!b || sample_stride < 1) return PM_ERR_INVALID;
  std::vector<uint2> sub;
  for (size_t i = 0; i < b->items_h.size(); i += (size_t)sample_stride) sub.push_back(b->items_h[i]);
  HIPCHK(b->items_tune_d.ensure(sizeof(uint2) * std::max<size_t>(sub.size(), 1)));
  if (!sub.empty())
    HIPCHK(hipMemcpy(b->items_tune_d.p, sub.data(), sizeof(uint2) * sub.size(), hipMemcpyHostToDevice));
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < 4; ++i)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  for (int i = 0; i < 4; ++i) HIPCHK(hipEventCreate(&ev[i]));
  // objective: kd-build + sampled render (the emission does not depend on the tuning)
  auto trial = [&](double* ms) -> int {
    HIPCHK(hipEventRecord(ev[0], b->stream));
    HIPCHK(enqueue_frame(b, b->items_tune_d.as<uint2>(), (uint32_t)sub.size(), ev[1], ev[2]));
    HIPCHK(hipEventRecord(ev[3], b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, ev[1], ev[3]));
    *ms = t;
    return PM_OK;
  };
  // staged sweep: heap keys (width x slots), then leaf size, then workgroup size
  pm_batch_tuning best = b->tune;
  double best_ms = -1.0;
  auto consider = [&](const pm_batch_tuning& t) -> int {
    if (t.heap_slots < b->kmax) return PM_OK;
    int r = apply_tuning(b, t, false);
    if (r == PM_ERR_INVALID) return PM_OK;  // a candidate this device cannot run: skipped
    if (r != PM_OK) return r;
    double ms = 0.0, ms2 = 0.0;
    if ((r = trial(&ms)) != PM_OK) return r;
    if ((r = trial(&ms2)) != PM_OK) return r;
    ms = std::min(ms, ms2);
    if (best_ms < 0.0 || ms < best_ms) {
      best_ms = ms;
      best = t;
    }
    return PM_OK;
  };
  const int keyb[3] = {8, 4, 2}, slots[2] = {32, 64}, leaves[4] = {4, 8, 16, 32}, blocks[3] = {64, 128, 256};
  const int tops[4] = {0, 3, 5, 7};
  const pm_batch_tuning start = b->tune;
  for (int kb : keyb)
    for (int sl : slots) {
      pm_batch_tuning t = start;
      t.heap_key_bytes = kb;
      t.heap_slots = sl;
      t.gather_mode = 0;
      t.top_levels = 0;
      int r = consider(t);
      if (r != PM_OK) return r;
    }
  {  // gather mode 7 (2-byte keys, 64 slots of LDS for its counters)
    pm_batch_tuning t = start;
    t.heap_key_bytes = 2;
    t.heap_slots = 64;
    t.gather_mode = 7;
    t.top_levels = 0;
    int r = consider(t);
    if (r != PM_OK) return r;
  }
  const pm_batch_tuning s1 = best;
  for (int L : leaves) {
    if (L == s1.leaf_size) continue;
    pm_batch_tuning t = s1;
    t.leaf_size = L;
    int r = consider(t);
    if (r != PM_OK) return r;
  }
  const pm_batch_tuning s2 = best;
  for (int bt : blocks) {
    if (bt == s2.block_threads) continue;
    pm_batch_tuning t = s2;
    t.block_threads = bt;
    int r = consider(t);
    if (r != PM_OK) return r;
  }
  if (best.gather_mode == 7) {  // LDS tile: top-of-tree levels staged per wave (mode 7 reads them)
    const pm_batch_tuning s3 = best;
    for (int lv : tops) {
      if (lv == s3.top_levels) continue;
      pm_batch_tuning t = s3;
      t.top_levels = lv;
      int r = consider(t);
      if (r != PM_OK) return r;
    }
  }
  int r = apply_tuning(b, best, true);
  if (r != PM_OK) return r;
  if (best_out) *best_out = best;
  if (best_ms_out) *best_ms_out = best_ms;
  return PM_OK;
}

int pm_batch_submit(pm_batch* b) {
  if (!b) return PM_ERR_INVALID;
  if (!b->exec) {
    int r = capture(b);
    if (r != PM_OK) return r;
  }
  double ms = 0.0;
  int r = run_frame(b, &ms);
  if (r != PM_OK) return r;
  unsigned long long q[2];
  HIPCHK(hipMemcpy(q, b->counters.p, 16, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(b->tally.data(), b->tally9.p, sizeof(double) * 9 * b->ns, hipMemcpyDeviceToHost));
  pm_batch_stats& s = b->st;
  s.n_scenes = (uint32_t)b->ns;
  s.work_items = (uint32_t)b->items_h.size();
  s.photons = 0;
  for (int i = 0; i < b->ns; ++i) s.photons += b->cfg[i].n_emit;
  s.surface_records = b->S_tot;
  s.volume_records = b->V_tot;
  s.pixels = b->npix;
  s.surface_queries = q[0];
  s.volume_queries = q[1];
  s.ms_frame = ms;
  s.tuning = b->tune;
  return PM_OK;
}

int pm_batch_time_phases(pm_batch* b, double* ms_emit, double* ms_build, double* ms_render) {
  if (!b || !ms_emit || !ms_build || !ms_render) return PM_ERR_INVALID;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < 4; ++i)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  for (int i = 0; i < 4; ++i) HIPCHK(hipEventCreate(&ev[i]));
  // the graph's kernels, launched one by one on the batch stream
  HIPCHK(hipEventRecord(ev[0], b->stream));
  HIPCHK(enqueue_frame(b, b->items_d.as<uint2>(), (uint32_t)b->items_h.size(), ev[1], ev[2]));
  HIPCHK(hipEventRecord(ev[3], b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  float t[3] = {0.0f, 0.0f, 0.0f};
  for (int i = 0; i < 3; ++i) HIPCHK(hipEventElapsedTime(&t[i], ev[i], ev[i + 1]));
  *ms_emit = t[0];
  *ms_build = t[1];
  *ms_render = t[2];
  return PM_OK;
}

int pm_batch_image(pm_batch* b, int scene, float* rgb_dev, pm_uchar4* out_dev) {
  if (!scene_ok(b, scene)) return PM_ERR_INVALID;
  const pm_config& c = b->cfg[scene];
  const size_t np = (size_t)c.width * (size_t)c.height;
  const float* src = b->rgb.as<float>() + 3 * b->img_off[scene];
  if (rgb_dev) HIPCHK(hipMemcpyAsync(rgb_dev, src, 12 * np, hipMemcpyDeviceToDevice, b->stream));
  if (out_dev) HIPCHK(launch_tonemap(src, np, c.exposure, (uchar4*)out_dev, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PM_OK;
}

int pm_batch_scene_stats(pm_batch* b, int scene, pm_stats* out) {
  if (!scene_ok(b, scene) || !out) return PM_ERR_INVALID;
  const pm_config& c = b->cfg[scene];
  memset(out, 0, sizeof(*out));
  out->n_surface = b->nrec_s[scene];
  out->n_volume = b->nrec_v[scene];
  out->n_emit_surface = out->n_emit_volume = out->n_emit_total = c.n_emit;
  const double* t = b->tally.data() + 9ull * scene;
  for (int k = 0; k < 3; ++k) {
    const double w = (double)c.light_power[k] / (double)c.n_emit;
    out->emitted[k] = (double)c.light_power[k];
    out->absorbed[k] = t[k] * w;
    out->escaped[k] = t[3 + k] * w;
    out->rr_gain[k] = t[6 + k] * w;
  }
  return PM_OK;
}

int pm_batch_stats_get(pm_batch* b, pm_batch_stats* out) {
  if (!b || !out) return PM_ERR_INVALID;
  *out = b->st;
  out->tuning = b->tune;
  return PM_OK;
}

int pm_batch_get_records(pm_batch* b, int scene, int map, float* pos, float* power, float* dir, uint32_t n) {
  if (!scene_ok(b, scene) || !pos || !power || !dir) return PM_ERR_INVALID;
  DevMap m = scene_map(b, scene, map);
  if (n != m.n) return PM_ERR_INVALID;
  return records_to_host(m, pos, power, dir);
}

int pm_batch_get_tree(pm_batch* b, int scene, int map, uint32_t* perm, float* split, uint32_t* axis, uint32_t n,
                      uint32_t ni) {
  if (!scene_ok(b, scene) || !perm) return PM_ERR_INVALID;
  DevMap m = scene_map(b, scene, map);
  if (n != m.n || ni != (m.nl > 0 ? m.nl - 1 : 0)) return PM_ERR_INVALID;
  return tree_to_host(m, perm, split, axis);
}

}  // extern "C"
