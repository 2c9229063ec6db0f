// batch.cpp -- batched execution of uneven transforms on one GPU.
//
// A batch of transforms of different sizes is spread over "lanes": HIP
// streams at the default priority, each with its own workspace slab, forked
// from the caller's stream by an event and joined back into it.  Items are
// assigned longest-first to the least-loaded lane (LPT on a cost model of
// contraction terms plus grid samples), so that the small transforms of a
// batch fill the machine around the large ones instead of running one after
// another.  Everything a batch needs (plans, Delta tables, DFT plans, index
// tables, lane slabs) is built before the first launch, so a batch can be
// captured into a HIP graph and replayed with no host work per item.
//
// A captured graph owns what its kernels address besides the caller's
// buffers: its own lane slabs (freed when the graph is destroyed, never
// shared with eager batches or other graphs) and a pin on every Delta table
// it reads (pinned tables are never evicted).  Replays of one graph are
// ordered: each launch waits for the previous one, on whatever stream.
#include "batch.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "conv.hpp"
#include "transform.hpp"

namespace so3g {

namespace {

constexpr int kMaxLanes = 8;

struct Lane {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  void *ws = nullptr;
  size_t cap = 0;  // double2 elements
};

// Buffers of the persistent mode: per-item workspaces (one arena), the
// device item tables of the grouped contraction launches and their queues.
struct GroupBufs {
  char *arena = nullptr;
  size_t arena_cap = 0;        // bytes
  CItem *itab = nullptr;
  size_t itab_cap = 0;         // entries
  unsigned *queues = nullptr;  // 8 per grouped launch
  size_t queue_cap = 0;        // launches
};

struct DevLanes {
  hipEvent_t fork = nullptr;
  Lane lane[kMaxLanes];
  // persistent mode: its phases run on stream gs (so eager batches take
  // turns on the shared GroupBufs); item tables go up from pinned memory
  hipStream_t gs = nullptr;
  hipEvent_t gfork = nullptr, gdone = nullptr, up_done = nullptr;
  bool up_pending = false;
  GroupBufs gb;
  CItem *host_itab = nullptr;
  size_t host_cap = 0;
};

enum Mode { kLanes = SO3X_BATCH_LANES, kPersistent = SO3X_BATCH_PERSISTENT };

std::mutex g_mu;
std::map<int, DevLanes> g_dev;
std::map<std::string, std::pair<int, int>> g_tuned;  // batch shape -> (mode, lane count)
struct GraphRec {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  std::vector<void *> slabs;                // lane workspaces the graph's kernels address
  std::vector<std::pair<int, int>> pins;    // (table L, dl_method) of the pinned Delta tables
  GroupBufs gb;                             // persistent mode: the graph's own buffers
  hipEvent_t done = nullptr;                // end of the last replay
  bool launched = false;
};
std::map<long long, GraphRec> g_graphs;
std::map<long long, hipEvent_t> g_tickets;
long long g_next_id = 1;

Method meth(int m) { return m == SO3X_DIRECT ? kDirect : kViaSsht; }

double item_cost(const so3x_item_t &it) {
  const so3_parameters_t &p = it.parameters;
  const double L = p.L, ns = p.reality ? p.N : 2.0 * p.N - 1.0;
  const double an = std::min<double>(p.N, p.L);
  // Delta-product terms: sum over n of sum_{l >= |n|} (2l+1)(l+1) ~ 2/3 (L^3 - |n|^3)
  const double terms = (ns * L * L * L - (p.reality ? 1.0 : 2.0) * an * an * an * an / 4.0) * 2.0 / 3.0;
  const double grid = ns * 2.0 * L * L;  // samples per DFT stage
  double c = terms + 60.0 * grid;
  if (it.kind == SO3X_ITEM_PAIR) c *= 2.0;
  return c;
}

std::string shape_key(const so3x_item_t *items, int count) {
  std::vector<std::string> v;
  for (int i = 0; i < count; ++i) {
    const so3_parameters_t &p = items[i].parameters;
    char b[160];
    snprintf(b, sizeof b, "%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d", items[i].kind, items[i].method, p.reality,
             p.L0, p.L, p.N, (int)p.sampling_scheme, (int)p.n_order, (int)p.storage, (int)p.n_mode,
             (int)p.dl_method, p.steerable, 0);
    v.emplace_back(b);
  }
  std::sort(v.begin(), v.end());
  std::string k;
  for (auto &s : v) k += s + ";";
  return k;
}

DevLanes &dev_lanes() {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  DevLanes &d = g_dev[dev];
  if (!d.fork) SO3_HIP(hipEventCreateWithFlags(&d.fork, hipEventDisableTiming));
  return d;
}

// plans, tables and workspace sizes of every item (allocates what is missing)
std::vector<size_t> prepare_items(const so3x_item_t *items, int count) {
  std::vector<size_t> ws(count);
  for (int i = 0; i < count; ++i) {
    const so3x_item_t &it = items[i];
    if (it.kind < SO3X_ITEM_INVERSE || it.kind > SO3X_ITEM_S2TOSO3_INVERSE) throw HipError("invalid item kind");
    if (it.method != SO3X_VIA_SSHT && it.method != SO3X_DIRECT) throw HipError("invalid item method");
    ws[i] = workspace_elems(&it.parameters, meth(it.method));
    if (it.kind == SO3X_ITEM_S2TOSO3_INVERSE) {
      if (it.parameters.steerable) throw HipError("s2toso3 needs steerable = 0");
      cached_bases(&it.parameters);
    }
  }
  return ws;
}

void run_item(const so3x_item_t &it, void *ws, hipStream_t s) {
  const so3_parameters_t *p = &it.parameters;
  const Method m = meth(it.method);
  switch (it.kind) {
    case SO3X_ITEM_INVERSE:
      enqueue_inverse(p, m, it.flmn_in, it.f, ws, s);
      break;
    case SO3X_ITEM_FORWARD:
      enqueue_forward(p, m, it.f, it.flmn_out, ws, s);
      break;
    case SO3X_ITEM_PAIR:
      enqueue_inverse(p, m, it.flmn_in, it.f, ws, s);
      enqueue_forward(p, m, it.f, it.flmn_out, ws, s);
      break;
    default:  // SO3X_ITEM_S2TOSO3_INVERSE
      s2toso3(p, it.flmn_in, it.glm, it.flmn_out, s);
      enqueue_inverse(p, m, it.flmn_out, it.f, ws, s);
      break;
  }
}

// Lane assignment of a batch: longest processing time first onto the
// least-loaded lane, and the workspace each lane needs.
struct LanePlan {
  int lanes = 1;
  std::vector<int> order, lane_of;
  std::vector<size_t> need;  // double2 elements per lane
};

LanePlan plan_lanes(const so3x_item_t *items, int count, int lanes, const std::vector<size_t> &ws) {
  LanePlan P;
  if (lanes <= 0) {
    auto it = g_tuned.find(shape_key(items, count));
    lanes = it != g_tuned.end() ? it->second.second : 4;
  }
  P.lanes = std::max(1, std::min({lanes, kMaxLanes, count}));
  P.order.resize(count);
  std::iota(P.order.begin(), P.order.end(), 0);
  std::vector<double> cost(count);
  for (int i = 0; i < count; ++i) cost[i] = item_cost(items[i]);
  std::stable_sort(P.order.begin(), P.order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  std::vector<double> load(P.lanes, 0.0);
  P.lane_of.resize(count);
  P.need.assign(P.lanes, 0);
  for (int i : P.order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    P.lane_of[i] = l;
    load[l] += cost[i];
    P.need[l] = std::max(P.need[l], ws[i]);
  }
  return P;
}

// the batch on `lanes` lanes, caller holds g_mu.  slabs == nullptr: the
// device's shared lane slabs (grown as needed); otherwise one caller-owned
// slab per lane of at least the planned size (a graph's own).
void enqueue_locked(const so3x_item_t *items, int count, int lanes, hipStream_t st,
                    const std::vector<void *> *slabs = nullptr) {
  if (count <= 0) return;
  if (!items) throw HipError("items pointer is NULL");
  const std::vector<size_t> ws = prepare_items(items, count);
  const LanePlan P = plan_lanes(items, count, lanes, ws);
  DevLanes &d = dev_lanes();
  std::vector<void *> use(P.lanes, nullptr);
  for (int l = 0; l < P.lanes; ++l) {
    Lane &ln = d.lane[l];
    if (!ln.s) {
      SO3_HIP(hipStreamCreateWithFlags(&ln.s, hipStreamNonBlocking));
      SO3_HIP(hipEventCreateWithFlags(&ln.done, hipEventDisableTiming));
    }
    if (slabs) {
      if ((int)slabs->size() < P.lanes) throw HipError("graph workspace does not match its batch");
      use[l] = (*slabs)[l];
      continue;
    }
    if (ln.cap < P.need[l]) {
      SO3_HIP(hipStreamSynchronize(ln.s));
      if (ln.ws) SO3_HIP(hipFree(ln.ws));
      ln.ws = nullptr;
      ln.cap = 0;
      SO3_HIP(hipMalloc(&ln.ws, sizeof(double2) * P.need[l]));
      ln.cap = P.need[l];
    }
    use[l] = ln.ws;
  }
  SO3_HIP(hipEventRecord(d.fork, st));
  for (int l = 0; l < P.lanes; ++l) SO3_HIP(hipStreamWaitEvent(d.lane[l].s, d.fork, 0));
  for (int i : P.order) run_item(items[i], use[P.lane_of[i]], d.lane[P.lane_of[i]].s);
  for (int l = 0; l < P.lanes; ++l) {
    SO3_HIP(hipEventRecord(d.lane[l].done, d.lane[l].s));
    SO3_HIP(hipStreamWaitEvent(st, d.lane[l].done, 0));
  }
}

// ------------------------------------------------------------ persistent
// The persistent mode runs a batch in three phases:
//   1. on stream gs: the s2toso3 products, then the inverse contractions of
//      every item in one grouped launch per signal type (complex, real):
//      a persistent grid pulling (item, block) work from device queues;
//   2. on the lanes (longest first): each item's stages between its
//      contractions (inverse beta, alpha, gamma; forward gamma, alpha, beta);
//   3. on gs: the forward contractions of every item, grouped likewise.
// Each item has its own workspace (its X1 / G' live across the phases), so
// the items' workspaces are summed; a batch whose sum exceeds kGroupBudget
// runs as consecutive groups within it.
constexpr size_t kGroupBudget = (size_t)64 << 30;

struct PGroup {
  std::vector<int> items;            // indices into the batch, lane order
  std::vector<size_t> ws_off;        // byte offset of each item's workspace
  size_t bytes = 0;
  int tab[4] = {0, 0, 0, 0};         // first entry of (inv complex, inv real, fwd complex, fwd real)
  int cnt[4] = {0, 0, 0, 0};
  unsigned blocks[4] = {0, 0, 0, 0};
};

struct PPlan {
  LanePlan lp;
  std::vector<PGroup> groups;
  std::vector<CItem> itab;           // host copy, slab addresses relative to arena 0
  size_t arena = 0;                  // bytes: the largest group
  size_t launches = 0;               // grouped launches (4 per group)
};

bool has_inverse(int kind) { return kind != SO3X_ITEM_FORWARD; }
bool has_forward(int kind) { return kind == SO3X_ITEM_FORWARD || kind == SO3X_ITEM_PAIR; }

// plan the phases; CItem addresses are filled for an arena at `arena_base`
PPlan plan_persistent(const so3x_item_t *items, int count, int lanes, char *arena_base) {
  PPlan P;
  const std::vector<size_t> ws = prepare_items(items, count);
  P.lp = plan_lanes(items, count, lanes, ws);
  PGroup cur;
  for (int i : P.lp.order) {
    const size_t b = round_up(sizeof(double2) * std::max<size_t>(ws[i], 1), 256);
    if (!cur.items.empty() && cur.bytes + b > kGroupBudget) {
      P.groups.push_back(cur);
      cur = PGroup{};
    }
    cur.items.push_back(i);
    cur.ws_off.push_back(cur.bytes);
    cur.bytes += b;
  }
  if (!cur.items.empty()) P.groups.push_back(cur);
  for (PGroup &g : P.groups) {
    P.arena = std::max(P.arena, g.bytes);
    std::vector<CItem> part[4];
    for (size_t k = 0; k < g.items.size(); ++k) {
      const so3x_item_t &it = items[g.items[k]];
      const so3_parameters_t *p = &it.parameters;
      const Method m = meth(it.method);
      void *slab = (void *)((uintptr_t)arena_base + g.ws_off[k]);
      const int r = p->reality ? 1 : 0;
      CItem c;
      if (has_inverse(it.kind)) {
        const void *in = it.kind == SO3X_ITEM_S2TOSO3_INVERSE ? it.flmn_out : it.flmn_in;
        contract_item(p, m, true, in, nullptr, slab, &c);
        part[r].push_back(c);
      }
      if (has_forward(it.kind)) {
        contract_item(p, m, false, nullptr, it.flmn_out, slab, &c);
        part[2 + r].push_back(c);
      }
    }
    for (int q = 0; q < 4; ++q) {
      g.tab[q] = (int)P.itab.size();
      g.cnt[q] = (int)part[q].size();
      unsigned off = 0;
      for (CItem &c : part[q]) {
        c.off = off;
        off += c.nblk;
        P.itab.push_back(c);
      }
      g.blocks[q] = off;
    }
    P.launches += 4;
  }
  return P;
}

// grow the shared buffers of the device (eager batches), ordered on gs
void grow_group_bufs(DevLanes &d, const PPlan &P) {
  GroupBufs &b = d.gb;
  if (b.arena_cap < P.arena || b.itab_cap < P.itab.size() || b.queue_cap < P.launches) {
    SO3_HIP(hipStreamSynchronize(d.gs));
    if (b.arena_cap < P.arena) {
      if (b.arena) SO3_HIP(hipFree(b.arena));
      b.arena = nullptr;
      b.arena_cap = 0;
      SO3_HIP(hipMalloc(&b.arena, P.arena));
      b.arena_cap = P.arena;
    }
    if (b.itab_cap < P.itab.size()) {
      if (b.itab) SO3_HIP(hipFree(b.itab));
      b.itab = nullptr;
      b.itab_cap = 0;
      SO3_HIP(hipMalloc(&b.itab, sizeof(CItem) * P.itab.size()));
      b.itab_cap = P.itab.size();
    }
    if (b.queue_cap < P.launches) {
      if (b.queues) SO3_HIP(hipFree(b.queues));
      b.queues = nullptr;
      b.queue_cap = 0;
      SO3_HIP(hipMalloc(&b.queues, sizeof(unsigned) * 8 * P.launches));
      b.queue_cap = P.launches;
    }
  }
}

// the three phases of every group on gs and the lanes (records kernels and
// event waits only, so it can be captured)
void launch_persistent(const so3x_item_t *items, const PPlan &P, const GroupBufs &b, DevLanes &d,
                       hipStream_t gs) {
  for (size_t gi = 0; gi < P.groups.size(); ++gi) {
    const PGroup &g = P.groups[gi];
    unsigned *q = b.queues + 8 * 4 * gi;
    for (int i : g.items)
      if (items[i].kind == SO3X_ITEM_S2TOSO3_INVERSE)
        s2toso3(&items[i].parameters, items[i].flmn_in, items[i].glm, items[i].flmn_out, gs);
    for (int k = 0; k < 2; ++k)
      launch_contract_grouped(b.itab + g.tab[k], g.cnt[k], g.blocks[k], k == 0 ? 2 : 1, +1, q + 8 * k, gs);
    SO3_HIP(hipEventRecord(d.gfork, gs));
    for (int l = 0; l < P.lp.lanes; ++l) SO3_HIP(hipStreamWaitEvent(d.lane[l].s, d.gfork, 0));
    for (size_t k = 0; k < g.items.size(); ++k) {
      const so3x_item_t &it = items[g.items[k]];
      hipStream_t ls = d.lane[P.lp.lane_of[g.items[k]]].s;
      void *slab = b.arena + g.ws_off[k];
      const Method m = meth(it.method);
      if (has_inverse(it.kind)) enqueue_inverse_post(&it.parameters, m, it.f, slab, ls);
      if (has_forward(it.kind)) enqueue_forward_pre(&it.parameters, m, it.f, it.flmn_out, slab, ls);
    }
    for (int l = 0; l < P.lp.lanes; ++l) {
      SO3_HIP(hipEventRecord(d.lane[l].done, d.lane[l].s));
      SO3_HIP(hipStreamWaitEvent(gs, d.lane[l].done, 0));
    }
    for (int k = 2; k < 4; ++k)
      launch_contract_grouped(b.itab + g.tab[k], g.cnt[k], g.blocks[k], k == 2 ? 2 : 1, -1, q + 8 * k, gs);
  }
}

void ensure_streams(DevLanes &d, int lanes) {
  for (int l = 0; l < lanes; ++l) {
    Lane &ln = d.lane[l];
    if (!ln.s) {
      SO3_HIP(hipStreamCreateWithFlags(&ln.s, hipStreamNonBlocking));
      SO3_HIP(hipEventCreateWithFlags(&ln.done, hipEventDisableTiming));
    }
  }
  if (!d.gs) {
    SO3_HIP(hipStreamCreateWithFlags(&d.gs, hipStreamNonBlocking));
    SO3_HIP(hipEventCreateWithFlags(&d.gfork, hipEventDisableTiming));
    SO3_HIP(hipEventCreateWithFlags(&d.gdone, hipEventDisableTiming));
    SO3_HIP(hipEventCreateWithFlags(&d.up_done, hipEventDisableTiming));
  }
}

// an eager persistent batch on the device's shared buffers
void enqueue_persistent_locked(const so3x_item_t *items, int count, int lanes, hipStream_t st) {
  if (count <= 0) return;
  if (!items) throw HipError("items pointer is NULL");
  DevLanes &d = dev_lanes();
  // plan once to size the buffers, then again with the arena's address
  PPlan P = plan_persistent(items, count, lanes, nullptr);
  ensure_streams(d, P.lp.lanes);
  grow_group_bufs(d, P);
  P = plan_persistent(items, count, lanes, d.gb.arena);
  // item tables: host staging (pinned) -> device, in order on gs
  if (d.up_pending) SO3_HIP(hipEventSynchronize(d.up_done));
  if (d.host_cap < P.itab.size()) {
    if (d.host_itab) SO3_HIP(hipHostFree(d.host_itab));
    d.host_itab = nullptr;
    d.host_cap = 0;
    SO3_HIP(hipHostMalloc((void **)&d.host_itab, sizeof(CItem) * P.itab.size(), hipHostMallocDefault));
    d.host_cap = P.itab.size();
  }
  std::copy(P.itab.begin(), P.itab.end(), d.host_itab);
  SO3_HIP(hipEventRecord(d.fork, st));
  SO3_HIP(hipStreamWaitEvent(d.gs, d.fork, 0));
  SO3_HIP(hipMemcpyAsync(d.gb.itab, d.host_itab, sizeof(CItem) * P.itab.size(), hipMemcpyHostToDevice, d.gs));
  SO3_HIP(hipEventRecord(d.up_done, d.gs));
  d.up_pending = true;
  launch_persistent(items, P, d.gb, d, d.gs);
  SO3_HIP(hipEventRecord(d.gdone, d.gs));
  SO3_HIP(hipStreamWaitEvent(st, d.gdone, 0));
}

void free_group_bufs(GroupBufs &b) {
  if (b.arena) (void)hipFree(b.arena);
  if (b.itab) (void)hipFree(b.itab);
  if (b.queues) (void)hipFree(b.queues);
  b = GroupBufs{};
}

void free_graph(GraphRec &rec) {
  if (rec.done) {
    (void)hipEventSynchronize(rec.done);
    (void)hipEventDestroy(rec.done);
  }
  if (rec.x) (void)hipGraphExecDestroy(rec.x);
  if (rec.g) (void)hipGraphDestroy(rec.g);
  for (void *p : rec.slabs) (void)hipFree(p);
  free_group_bufs(rec.gb);
  for (auto &pn : rec.pins) dl_tables_unpin(pn.first, pn.second);
  rec = GraphRec{};
}

int resolve_mode(const so3x_item_t *items, int count, int &lanes, int mode) {
  if (mode == kLanes || mode == kPersistent) return mode;
  if (lanes > 0) return kLanes;  // an explicit lane count without a mode: lanes
  auto it = g_tuned.find(shape_key(items, count));
  if (it == g_tuned.end()) return kLanes;
  lanes = it->second.second;
  return it->second.first;
}

}  // namespace

void batch_enqueue(const so3x_item_t *items, int count, int lanes, int mode, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  mode = resolve_mode(items, count, lanes, mode);
  if (mode == kPersistent) enqueue_persistent_locked(items, count, lanes, st);
  else enqueue_locked(items, count, lanes, st);
}

void batch_choice(const so3x_item_t *items, int count, int *mode, int *lanes) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_tuned.find(shape_key(items, count));
  if (mode) *mode = it == g_tuned.end() ? kLanes : it->second.first;
  if (lanes) *lanes = it == g_tuned.end() ? 4 : it->second.second;
}

int batch_tune(const so3x_item_t *items, int count, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (count <= 0) return 1;
  hipEvent_t a, b;
  SO3_HIP(hipEventCreate(&a));
  SO3_HIP(hipEventCreate(&b));
  int best = 1, best_mode = kLanes;
  float best_ms = 1e30f;
  for (int mode : {(int)kLanes, (int)kPersistent}) {
    for (int lanes : {1, 2, 4, 8}) {
      if (lanes > count && lanes > 1) break;
      auto run = [&] {
        if (mode == kPersistent) enqueue_persistent_locked(items, count, lanes, st);
        else enqueue_locked(items, count, lanes, st);
      };
      run();  // warm-up (allocates workspaces)
      SO3_HIP(hipEventRecord(a, st));
      for (int r = 0; r < 3; ++r) run();
      SO3_HIP(hipEventRecord(b, st));
      SO3_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      SO3_HIP(hipEventElapsedTime(&ms, a, b));
      if (ms < best_ms) best_ms = ms, best = lanes, best_mode = mode;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  g_tuned[shape_key(items, count)] = std::make_pair(best_mode, best);
  return best;
}

long long graph_create(const so3x_item_t *items, int count, int lanes, int mode) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (count <= 0 || !items) throw HipError("a graph needs at least one item");
  mode = resolve_mode(items, count, lanes, mode);
  hipStream_t cap = nullptr;
  SO3_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  GraphRec rec;
  try {
    // one eager run builds every plan, table and DFT plan and sets the
    // kernels' LDS attributes, so the capture records launches only
    if (mode == kPersistent) enqueue_persistent_locked(items, count, lanes, cap);
    else enqueue_locked(items, count, lanes, cap);
    SO3_HIP(hipStreamSynchronize(cap));
    // the graph's own workspaces and pins on its Delta tables
    const std::vector<size_t> ws = prepare_items(items, count);
    const LanePlan P = plan_lanes(items, count, lanes, ws);
    PPlan PP;
    if (mode == kPersistent) {
      PP = plan_persistent(items, count, lanes, nullptr);
      GroupBufs &b = rec.gb;
      SO3_HIP(hipMalloc(&b.arena, std::max<size_t>(PP.arena, 256)));
      b.arena_cap = PP.arena;
      SO3_HIP(hipMalloc(&b.itab, sizeof(CItem) * std::max<size_t>(PP.itab.size(), 1)));
      b.itab_cap = PP.itab.size();
      SO3_HIP(hipMalloc(&b.queues, sizeof(unsigned) * 8 * std::max<size_t>(PP.launches, 1)));
      b.queue_cap = PP.launches;
      PP = plan_persistent(items, count, lanes, b.arena);
      SO3_HIP(hipMemcpy(b.itab, PP.itab.data(), sizeof(CItem) * PP.itab.size(), hipMemcpyHostToDevice));
    } else {
      for (int l = 0; l < P.lanes; ++l) {
        void *p = nullptr;
        SO3_HIP(hipMalloc(&p, sizeof(double2) * std::max<size_t>(P.need[l], 1)));
        rec.slabs.push_back(p);
      }
    }
    // pin the table that serves each item (built by the eager run); the
    // capture then records exactly these tables
    for (int i = 0; i < count; ++i) {
      const int L = items[i].parameters.L, m = (int)items[i].parameters.dl_method;
      const std::pair<int, int> k(dl_tables(L, m).L, m);
      if (std::find(rec.pins.begin(), rec.pins.end(), k) != rec.pins.end()) continue;
      if (dl_tables_pin(L, m) != k.first) throw HipError("Delta table changed while pinning");
      rec.pins.push_back(k);
    }
    SO3_HIP(hipEventCreateWithFlags(&rec.done, hipEventDisableTiming));
    SO3_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed));
    try {
      if (mode == kPersistent) {
        // the phases on the capture stream itself (gs of the recording)
        DevLanes &d = dev_lanes();
        ensure_streams(d, PP.lp.lanes);
        launch_persistent(items, PP, rec.gb, d, cap);
      } else {
        enqueue_locked(items, count, P.lanes, cap, &rec.slabs);
      }
    } catch (...) {
      hipGraph_t dummy = nullptr;
      (void)hipStreamEndCapture(cap, &dummy);
      if (dummy) (void)hipGraphDestroy(dummy);
      throw;
    }
    SO3_HIP(hipStreamEndCapture(cap, &rec.g));
    SO3_HIP(hipGraphInstantiate(&rec.x, rec.g, nullptr, nullptr, 0));
  } catch (...) {
    free_graph(rec);
    (void)hipStreamDestroy(cap);
    throw;
  }
  SO3_HIP(hipStreamDestroy(cap));
  const long long id = g_next_id++;
  g_graphs[id] = rec;
  return id;
}

void graph_launch(long long graph, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_graphs.find(graph);
  if (it == g_graphs.end()) throw HipError("unknown graph handle");
  GraphRec &rec = it->second;
  // replays share the graph's slabs: each waits for the previous one
  if (rec.launched) SO3_HIP(hipStreamWaitEvent(st, rec.done, 0));
  SO3_HIP(hipGraphLaunch(rec.x, st));
  SO3_HIP(hipEventRecord(rec.done, st));
  rec.launched = true;
}

void graph_destroy(long long graph) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_graphs.find(graph);
  if (it == g_graphs.end()) throw HipError("unknown graph handle");
  free_graph(it->second);  // waits for the last replay first
  g_graphs.erase(it);
}

long long ticket_new(hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t e = nullptr;
  SO3_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  SO3_HIP(hipEventRecord(e, st));
  const long long id = g_next_id++;
  g_tickets[id] = e;
  return id;
}

int ticket_done(long long ticket) {
  hipEvent_t e = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tickets.find(ticket);
    if (it == g_tickets.end()) throw HipError("unknown ticket");
    e = it->second;
  }
  const hipError_t r = hipEventQuery(e);
  if (r == hipSuccess) return 1;
  if (r == hipErrorNotReady) return 0;
  SO3_HIP(r);
  return 0;
}

void ticket_wait(long long ticket) {
  hipEvent_t e = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tickets.find(ticket);
    if (it == g_tickets.end()) throw HipError("unknown ticket");
    e = it->second;
    g_tickets.erase(it);
  }
  const hipError_t r = hipEventSynchronize(e);
  (void)hipEventDestroy(e);
  SO3_HIP(r);
}

void release_batch() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_graphs) free_graph(kv.second);
  g_graphs.clear();
  for (auto &kv : g_tickets) (void)hipEventDestroy(kv.second);
  g_tickets.clear();
  for (auto &kv : g_dev) {
    if (kv.second.gs) (void)hipStreamSynchronize(kv.second.gs);
    free_group_bufs(kv.second.gb);
    if (kv.second.host_itab) (void)hipHostFree(kv.second.host_itab);
    if (kv.second.gfork) (void)hipEventDestroy(kv.second.gfork);
    if (kv.second.gdone) (void)hipEventDestroy(kv.second.gdone);
    if (kv.second.up_done) (void)hipEventDestroy(kv.second.up_done);
    if (kv.second.gs) (void)hipStreamDestroy(kv.second.gs);
    for (Lane &ln : kv.second.lane) {
      if (ln.s) (void)hipStreamSynchronize(ln.s);
      if (ln.ws) (void)hipFree(ln.ws);
      if (ln.done) (void)hipEventDestroy(ln.done);
      if (ln.s) (void)hipStreamDestroy(ln.s);
      ln = Lane{};
    }
    if (kv.second.fork) (void)hipEventDestroy(kv.second.fork);
  }
  g_dev.clear();
}

}  // namespace so3g
