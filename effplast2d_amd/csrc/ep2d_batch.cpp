// ep2d_batch.cpp -- batched multi-RVE engine (see ep2d_batch.h).
//
// Each admitted request owns a Solver (its fields, masks, reduction scratch)
// on the engine's stream.  An iteration of all admitted RVEs is one launch of
// k_fused_batch over the concatenated tile lists; a chunk is the run of
// launches up to the next residual check (the check interval is common, so
// all RVEs reach their checks together).  Between chunks the host checks the
// residuals (one synchronisation for all RVEs), post-processes finished load
// steps, retires finished jobs and admits queued requests.  An RVE's
// arithmetic does not depend on its neighbours in the launch, so its result
// is the same as on the single-RVE path.
#include "ep2d_batch.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

namespace ep2d {

namespace {
constexpr int kTileX = 62;  // fused tile width (Fused::TX)
}  // namespace

// Kernel shapes the tuner times: the fused tile kernel, the two-iteration
// wave pairs at two segment heights, the chains of three and four stages on
// one wave per stage, and the four-stage chains on two waves per stage (fewer
// halo columns per output, six waves per SIMD instead of seven).  Short
// segments give small batches enough workgroups; tall ones have fewer halo
// rows.  The wide chains (slower on the batches measured) and the three-
// stage chains on two waves per stage (six-wave workgroups spread unevenly
// over a CU's four SIMDs) stay fixed shapes.  Two timed rounds of every
// shape fit a 200-iteration chunk.
const Batch::Shape Batch::kCand[Batch::kNCand] = {
    {kTiles, 8},    {kPair, 14},    {kPair, 26},    {kTriple, 42},  {kTriple, 128},
    {kQuad, 64},    {kQuad, 128},   {kQuad2, 48},   {kQuad2, 64},   {kQuad2, 96},
    {kQuad2, 128}};

namespace {

bool valid_mode(int m) { return m == EP2D_MODE_FP64 || m == EP2D_MODE_SPLIT; }
}  // namespace

Batch::Batch(const ep2d_batch_options& opt) : opt_(opt) {
  if (!valid_mode(opt_.mode)) throw ArgError("batch mode must be EP2D_MODE_FP64 or EP2D_MODE_SPLIT");
  if (opt_.checkInterval <= 0) opt_.checkInterval = 10000;
  if (opt_.tileRows != 0 && !valid_tile_rows(opt_.tileRows))
    throw ArgError("tileRows must be 0, 6, 8, 14 or 30");
  if (opt_.marchRows < 0) throw ArgError("marchRows must be >= 0");
  if (opt_.persistRows != 0 && !valid_tile_rows(opt_.persistRows))
    throw ArgError("persistRows must be 0, 6, 8, 14 or 30");
  if (opt_.march2Rows < 0) throw ArgError("march2Rows must be >= 0");
  if (opt_.pairRows < 0) throw ArgError("pairRows must be >= 0");
  if (opt_.tripleRows < 0) throw ArgError("tripleRows must be >= 0");
  if (opt_.quadRows < 0) throw ArgError("quadRows must be >= 0");
  if (opt_.wide3Rows < 0) throw ArgError("wide3Rows must be >= 0");
  if (opt_.wide4Rows < 0) throw ArgError("wide4Rows must be >= 0");
  if (opt_.triple2Rows < 0) throw ArgError("triple2Rows must be >= 0");
  if (opt_.quad2Rows < 0) throw ArgError("quad2Rows must be >= 0");
  if ((opt_.wide3Rows != 0 || opt_.wide4Rows != 0) && opt_.mode != EP2D_MODE_SPLIT)
    throw ArgError("wide3Rows and wide4Rows (wide chains) need the split scheme");
  const int nfixed = (opt_.tileRows != 0) + (opt_.marchRows != 0) + (opt_.persistRows != 0) +
                     (opt_.march2Rows != 0) + (opt_.pairRows != 0) + (opt_.tripleRows != 0) +
                     (opt_.quadRows != 0) + (opt_.wide3Rows != 0) + (opt_.wide4Rows != 0) +
                     (opt_.triple2Rows != 0) + (opt_.quad2Rows != 0);
  if (nfixed > 1)
    throw ArgError("give at most one of tileRows, marchRows, persistRows, march2Rows, pairRows, "
                   "tripleRows, quadRows, wide3Rows, wide4Rows, triple2Rows, quad2Rows");
  fixed_ = nfixed == 1;
  if (opt_.marchRows) shape_ = {kMarch, opt_.marchRows};
  else if (opt_.tileRows) shape_ = {kTiles, opt_.tileRows};
  else if (opt_.persistRows) shape_ = {kPersist, opt_.persistRows};
  else if (opt_.march2Rows) shape_ = {kMarch2, opt_.march2Rows};
  else if (opt_.pairRows) shape_ = {kPair, opt_.pairRows};
  else if (opt_.tripleRows) shape_ = {kTriple, opt_.tripleRows};
  else if (opt_.quadRows) shape_ = {kQuad, opt_.quadRows};
  else if (opt_.wide3Rows) shape_ = {kWide3, opt_.wide3Rows};
  else if (opt_.wide4Rows) shape_ = {kWide4, opt_.wide4Rows};
  else if (opt_.triple2Rows) shape_ = {kTriple2, opt_.triple2Rows};
  else if (opt_.quad2Rows) shape_ = {kQuad2, opt_.quad2Rows};
  device_ = opt_.device;
  int ndev = 0;
  EP2D_HIP(hipGetDeviceCount(&ndev));
  if (device_ < 0 || device_ >= ndev) throw ArgError("device index out of range");
  EP2D_HIP(hipSetDevice(device_));
  EP2D_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  size_t freeB = 0, totalB = 0;
  EP2D_HIP(hipMemGetInfo(&freeB, &totalB));
  budget_ = opt_.memoryBudget ? opt_.memoryBudget : (size_t)(0.9 * (double)freeB);
  t0_ = std::chrono::steady_clock::now();
}

Batch::~Batch() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  active_.clear();
  dropGraphs();
  for (auto& e : evPending_) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  for (auto& e : evPool_) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  if (dRves_) (void)hipFree(dRves_);
  if (dTiles_) (void)hipFree(dTiles_);
  if (hRves_) (void)hipHostFree(hRves_);
  if (hTiles_) (void)hipHostFree(hTiles_);
  if (dJobs_) (void)hipFree(dJobs_);
  if (dRes_) (void)hipFree(dRes_);
  if (hJobs_) (void)hipHostFree(hJobs_);
  if (hRes_) (void)hipHostFree(hRes_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

double Batch::clock() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

long long Batch::submit(const ep2d_request& rq) {
  if (rq.nX < 8) throw ArgError("request: nX must be >= 8");
  if (rq.nTasks < 1 || rq.nIter < 1 || rq.nTimeSteps < 1)
    throw ArgError("request: nTasks, nIter and nTimeSteps must be >= 1");
  if (rq.checkInterval < 0) throw ArgError("request: checkInterval must be >= 0");
  if (fixed_ && multiIter(shape_.kind) && !multi_iter_fits(rq.nX, rq.nX))
    throw ArgError(
        "request: march2Rows, pairRows, tripleRows, quadRows and the wide chains need every field "
        "under 2 GiB "
        "(nX <= 16382)");
  queue_.push_back({nextId_, rq, clock()});
  st_.submitted++;
  return nextId_++;
}

// Device bytes one RVE holds (the Solver's allocations), to admit within
// budget.
size_t Batch::estimateBytes(const ep2d_request& rq) const {
  return Solver::plannedDeviceBytes(rq.nX, rq.nX, opt_.mode, rq.pa[EP2D_PA_NPORES]);
}

void Batch::admit() {
  while (!queue_.empty()) {
    if (opt_.maxActive > 0 && (int)active_.size() >= opt_.maxActive) return;
    Job& jb = queue_.front();
    const size_t need = estimateBytes(jb.rq);
    if (used_ + need > budget_ && !active_.empty()) return;  // wait for retirements
    // blocks held for reuse count against the budget too
    if (used_ + need + pool_.heldBytes() > budget_) pool_.trim();
    Slot sl;
    sl.id = jb.id;
    sl.tag = jb.rq.tag;
    sl.tSubmit = jb.tSubmit;
    sl.tAdmit = clock();
    ep2d_options o;
    std::memset(&o, 0, sizeof(o));
    o.nTasks = jb.rq.nTasks;
    o.nIter = jb.rq.nIter;
    o.eIter = jb.rq.eIter;
    o.checkInterval = jb.rq.checkInterval > 0 ? jb.rq.checkInterval : opt_.checkInterval;
    o.mode = opt_.mode;
    o.device = device_;
    o.nTimeSteps = jb.rq.nTimeSteps;
    o.useShearLoadType = jb.rq.useShearLoadType;
    std::memcpy(o.shearLoadType, jb.rq.shearLoadType, sizeof(o.shearLoadType));
    o.tileRows = 0;
    o.itersPerLaunch = 1;   // the engine launches its own kernels
    o.adaptiveCheck = opt_.adaptiveCheck;
    o.keepFp32Stress = opt_.keepFp32Stress;
    try {
      sl.s = std::make_unique<Solver>(jb.rq.pa, jb.rq.nX, jb.rq.nX, o, stream_, &pool_);
      sl.bytes = sl.s->deviceBytes();
      sl.s->jobBegin(jb.rq.nTasks > 3 ? 1 : 0, jb.rq.initLoadValue, jb.rq.loadType,
                     jb.rq.loadValue);
      sl.s->tsBegin();
    } catch (const StatusError& e) {
      fail(sl, e.status, e.what());
      queue_.pop_front();
      continue;
    } catch (const ArgError& e) {
      fail(sl, EP2D_ERR_ARG, e.what());
      queue_.pop_front();
      continue;
    } catch (const HipError& e) {
      fail(sl, EP2D_ERR_HIP, e.what());
      queue_.pop_front();
      continue;
    }
    used_ += sl.bytes;
    active_.push_back(std::move(sl));
    queue_.pop_front();
  }
}

void Batch::fail(Slot& sl, int status, const char* what) {
  ep2d_batch_result r;
  std::memset(&r, 0, sizeof(r));
  r.id = sl.id;
  r.tag = sl.tag;
  r.status = status;
  r.tSubmit = sl.tSubmit;
  r.tAdmit = sl.tAdmit;
  r.tDone = clock();
  std::snprintf(r.message, sizeof(r.message), "%s", what ? what : "");
  finished_.push_back(r);
  st_.failed++;
}

void Batch::finish(Slot& sl) {
  ep2d_batch_result r;
  std::memset(&r, 0, sizeof(r));
  r.id = sl.id;
  r.tag = sl.tag;
  r.status = EP2D_OK;
  sl.s->jobFinish(&r.result);
  st_.fp64Finishes += sl.s->escalations();
  r.tSubmit = sl.tSubmit;
  r.tAdmit = sl.tAdmit;
  r.tDone = clock();
  finished_.push_back(r);
  st_.finished++;
}

void Batch::ensureCapacity(size_t nRves, size_t nTiles) {
  if (nRves > capRves_) {
    const size_t cap = std::max<size_t>(nRves, 2 * capRves_);
    dropGraphs();
    if (dRves_) EP2D_HIP(hipFree(dRves_));
    if (hRves_) EP2D_HIP(hipHostFree(hRves_));
    dRves_ = nullptr;
    hRves_ = nullptr;
    EP2D_HIP(hipMalloc((void**)&dRves_, cap * sizeof(RveDev)));
    EP2D_HIP(hipHostMalloc((void**)&hRves_, cap * sizeof(RveDev), hipHostMallocDefault));
    capRves_ = cap;
  }
  if (nTiles > capTiles_) {
    const size_t cap = std::max<size_t>(nTiles, 2 * capTiles_);
    dropGraphs();
    if (dTiles_) EP2D_HIP(hipFree(dTiles_));
    if (hTiles_) EP2D_HIP(hipHostFree(hTiles_));
    dTiles_ = nullptr;
    hTiles_ = nullptr;
    EP2D_HIP(hipMalloc((void**)&dTiles_, cap * sizeof(int4)));
    EP2D_HIP(hipHostMalloc((void**)&hTiles_, cap * sizeof(int4), hipHostMallocDefault));
    capTiles_ = cap;
  }
}

// Room for n residual jobs (called between chunks, after the stream was
// synchronised, so no copy or kernel still uses the old buffers).
void Batch::ensureJobs(size_t n) {
  if (n <= capJobs_) return;
  const size_t cap = std::max<size_t>(n, 2 * capJobs_);
  if (dJobs_) EP2D_HIP(hipFree(dJobs_));
  if (dRes_) EP2D_HIP(hipFree(dRes_));
  if (hJobs_) EP2D_HIP(hipHostFree(hJobs_));
  if (hRes_) EP2D_HIP(hipHostFree(hRes_));
  dJobs_ = nullptr;
  dRes_ = nullptr;
  hJobs_ = nullptr;
  hRes_ = nullptr;
  capJobs_ = 0;
  EP2D_HIP(hipMalloc((void**)&dJobs_, cap * sizeof(MaxabsJob)));
  EP2D_HIP(hipMalloc((void**)&dRes_, 3 * cap * sizeof(unsigned long long)));
  EP2D_HIP(hipHostMalloc((void**)&hJobs_, cap * sizeof(MaxabsJob), hipHostMallocDefault));
  EP2D_HIP(hipHostMalloc((void**)&hRes_, 3 * cap * sizeof(unsigned long long),
                         hipHostMallocDefault));
  capJobs_ = cap;
}

void Batch::dropGraphs() {
  for (auto& g : graphs_) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  graphs_.clear();
}

// Work list of every admitted RVE for a kernel shape.  Tiles (fused tile
// kernels): ceil(nX/62) columns of 62 outputs by ceil(nY/rows) rows, entries
// (RVE index, i0, j0, unused).  Segments (marching kernels): strips of 60
// (one iteration), 56 (two), 52 (three) or 48 (four) output columns by runs
// of `rows` rows, entries (RVE index, s, j0, j1).  The RVE's current buffers are in its
// table entry, so the list stays valid while the RVEs iterate.
int Batch::buildWork(Shape sh, int4* out, bool esc) const {
  int t = 0;
  for (size_t k = 0; k < active_.size(); k++) {
    const Solver& s = *active_[k].s;
    // the RVEs finishing a time step in the FP64 arithmetic are in the
    // escalated lists only (esc, or the tiles of 8 rows of kEsc), the others
    // in every other list
    if (s.escalated() != (esc || sh.kind == kEsc)) continue;
    if (sh.kind == kMarch) {
      t += march_segments(s.nX(), s.nY(), sh.rows, (int)k, out ? out + t : nullptr);
      continue;
    }
    if (chain(sh.kind)) {
      t += chain_segments(ipl(sh.kind), s.nX(), s.nY(), sh.rows, (int)k, out ? out + t : nullptr,
                          stageWaves(sh.kind));
      continue;
    }
    if (wide(sh.kind)) {
      t += wide_segments(ipl(sh.kind), s.nX(), s.nY(), sh.rows, (int)k, out ? out + t : nullptr);
      continue;
    }
    if (multiIter(sh.kind)) {
      t += march2_segments(s.nX(), s.nY(), sh.rows, (int)k, out ? out + t : nullptr);
      continue;
    }
    const int rows = sh.kind == kEsc ? 8 : sh.rows;
    const int ntx = (s.nX() + kTileX - 1) / kTileX;
    const int nty = (s.nY() + rows - 1) / rows;
    for (int j = 0; j < nty; j++)
      for (int i = 0; i < ntx; i++) {
        if (out) out[t] = make_int4((int)k, i * kTileX, j * rows, 0);
        t++;
      }
  }
  if (chain(sh.kind)) {
    std::vector<int2> dims(active_.size());
    for (size_t k = 0; k < active_.size(); k++)
      dims[k] = make_int2(active_[k].s->nX(), active_[k].s->nY());
    const int res = chain_resident(ipl(sh.kind), stageWaves(sh.kind),
                                   opt_.mode == EP2D_MODE_SPLIT && !esc);
    t = chain_work_order(out, t, stageWaves(sh.kind), dims.data(), res, out);
  }
  return t;
}

// One launch of shape sh as the l-th launch of the chunk after d multi-
// iteration launches: U parity l, V parity d.  Two- and three-iteration
// shapes launch their kernel when dbl, else one marching iteration over the
// auxiliary list (V in place in the current buffer).  The other kernels
// update V in place in the table's buffer, so they are only enqueued when d
// is even.
void Batch::launchOne(Shape sh, const int4* work, int n, bool dbl, long l, long d, hipEvent_t e0,
                      hipEvent_t e1) {
  if (chain(sh.kind) && dbl)
    launch_chain(ipl(sh.kind), dRves_, work, n, opt_.mode, (int)(l & 1), (int)(d & 1), stream_,
                 e0, e1, stageWaves(sh.kind));
  else if (wide(sh.kind) && dbl)
    launch_wide(ipl(sh.kind), dRves_, work, n, (int)(l & 1), (int)(d & 1), stream_, e0, e1);
  else if (multiIter(sh.kind) && dbl)
    launch_march2(dRves_, work, n, opt_.mode, (int)(l & 1), (int)(d & 1), stream_, e0, e1,
                  sh.kind == kPair);
  else if (multiIter(sh.kind))
    launch_march(dRves_, dTiles_ + auxOff_, auxCnt_, opt_.mode, (int)(l & 1), (int)(d & 1),
                 stream_, e0, e1);
  else if (sh.kind == kMarch)
    launch_march(dRves_, work, n, opt_.mode, (int)(l & 1), (int)(d & 1), stream_, e0, e1);
  else if (sh.kind == kPersist)
    launch_fused_persist(dRves_, work, n, opt_.mode, sh.rows, (int)(l & 1), stream_, e0, e1);
  else
    launch_fused_batch(dRves_, work, n, opt_.mode, sh.rows, (int)(l & 1), stream_, e0, e1);
}

int Batch::candIndex(Shape sh) {
  for (int k = 0; k < kNCand; k++)
    if (kCand[k].kind == sh.kind && kCand[k].rows == sh.rows) return k;
  return 0;
}

// Split of `iters` iterations for a shape of n iterations per launch: the
// multi-iteration launches first, then 1..n single iterations, so the chunk's
// last launch reads U^{n-1} (the stress fields are recomputed from that
// buffer after a load step).
static void split_multi(long iters, int n, long& packs, long& singles) {
  singles = iters <= 0 ? 0 : (iters - 1) % n + 1;
  packs = (iters - singles) / n;
}

// `n` launches of the current shape from launch index l and double index d
// (pairs when dbl); advances l and d.
void Batch::enqueue(long n, bool dbl, long& l, long& d) {
  for (long k = 0; k < n; k++) {
    launchOne(shape_, dTiles_, nTiles_, dbl, l, d);
    l++;
    if (dbl) d++;
  }
}

// The shapes of the closing launches (rem3Off_, rem2Off_), and the chain
// shape of the RVEs finishing a time step in the FP64 arithmetic.
const Batch::Shape Batch::kClose3{Batch::kTriple, 42};
const Batch::Shape Batch::kClose2{Batch::kPair, 26};
const Batch::Shape Batch::kEscMain{Batch::kQuad2, 64};

// c iterations of the RVEs finishing a time step in the FP64 arithmetic (FP64
// stress and velocity arithmetic on the split state: FP64 U, FP32 V), with
// the U and V parities of a chunk of their own.  On the chain kernels (the
// four-stage chains on two waves, closed like every chunk, EffPlast2D.cu:428)
// when they all fit them, else c launches of the tile kernel (U flips each
// launch, V in place).
Batch::Seq Batch::runEsc(long c) {
  if (!escChains_) {
    for (long k = 0; k < c; k++)
      launch_fused_batch(dRves_, dTiles_ + escOff_, escCnt_, kModeSplitStress64, 8,
                         (int)(k & 1), stream_);
    return Seq{c, 0};
  }
  const int m = kModeSplitStress64;
  long packs = 0, singles = 0, l = 0, d = 0;
  split_multi(c, 4, packs, singles);
  for (long k = 0; k < packs; k++, l++, d++)
    launch_chain(4, dRves_, dTiles_ + escMainOff_, escMainCnt_, m, (int)(l & 1), (int)(d & 1),
                 stream_, nullptr, nullptr, 2);
  if (singles == 4) {
    launch_chain(3, dRves_, dTiles_ + escC3Off_, escC3Cnt_, m, (int)(l & 1), (int)(d & 1), stream_,
                 nullptr, nullptr, 1);
    l++, d++, singles = 1;
  } else if (singles == 3) {
    launch_march2(dRves_, dTiles_ + escC2Off_, escC2Cnt_, m, (int)(l & 1), (int)(d & 1), stream_,
                  nullptr, nullptr, true);
    l++, d++, singles = 1;
  }
  for (long k = 0; k < singles; k++, l++)
    launch_march(dRves_, dTiles_ + escAuxOff_, escAuxCnt_, m, (int)(l & 1), (int)(d & 1),
                 stream_);
  return Seq{l, d};
}

// The 1..n single iterations that close a chunk of an n-iteration shape:
// the last one is a single-iteration launch (it reads U^{n-1}); four of them
// run as a three-iteration chain launch and the single, three as a wave-pair
// launch and the single (the same results, in a third of the time).
void Batch::closing(long singles, long& l, long& d) {
  if (singles == 4 && rem3Cnt_ > 0) {
    launchOne(kClose3, dTiles_ + rem3Off_, rem3Cnt_, true, l, d);
    l++;
    d++;
    singles = 1;
  } else if (singles == 3 && rem2Cnt_ > 0) {
    launchOne(kClose2, dTiles_ + rem2Off_, rem2Cnt_, true, l, d);
    l++;
    d++;
    singles = 1;
  }
  enqueue(singles, false, l, d);
}

// Graph of `len` launches of the current shape (multi-iteration launches for
// a multi-iteration shape) from launch and double index 0 over the current
// tables, captured once per length while the tables stay the same.
hipGraphExec_t Batch::graphFor(long len) {
  for (auto& g : graphs_)
    if (g.len == len) {
      g.lastUse = st_.chunks;
      return g.exec;
    }
  if ((int)graphs_.size() >= kMaxGraphs) {
    auto lru = graphs_.begin();
    for (auto it = graphs_.begin(); it != graphs_.end(); ++it)
      if (it->lastUse < lru->lastUse) lru = it;
    (void)hipGraphExecDestroy(lru->exec);
    (void)hipGraphDestroy(lru->graph);
    graphs_.erase(lru);
  }
  GraphEntry g;
  g.len = len;
  g.lastUse = st_.chunks;
  long l = 0, d = 0;
  EP2D_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  enqueue(len, multiIter(shape_.kind), l, d);
  EP2D_HIP(hipStreamEndCapture(stream_, &g.graph));
  EP2D_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
  graphs_.push_back(g);
  return g.exec;
}

// c iterations of every admitted RVE.  Launch l reads U buffer cur_k ^ (l&1)
// of RVE k (cur_k in its table entry); multi-iteration launch number d reads
// V buffer d&1 (the table's Vx/Vy are the current ones) and writes the other.
// The shape does not change any result, so a tuning chunk spends its first
// launches timing each candidate (multi-iteration candidates in pairs of
// launches, so V is back in its buffer before the next candidate).  Returns
// the launch counts.
Batch::Seq Batch::launchChunk(long c) {
  std::vector<long long> ids;
  ids.reserve(active_.size() + 3);
  double bytes = 0.0, cells = 0.0;
  bool multiOk = true;  // every field under the multi-iteration kernels' bound
  for (auto& sl : active_) {
    // an RVE finishing a time step in the FP64 arithmetic is stepped on its
    // own (Batch::step): not in the work lists, the statistics or the tuning
    ids.push_back(2 * sl.id + (sl.s->escalated() ? 1 : 0));
    if (sl.s->escalated()) continue;
    const double n = sl.s->nX();
    bytes += (opt_.mode == EP2D_MODE_SPLIT ? 48.0 : 64.0) * (n + 1) * n;
    cells += n * n;
    multiOk = multiOk && multi_iter_fits(sl.s->nX(), sl.s->nX());
  }
  if (!fixed_ && !tuneDue_ && (cells > 1.25 * tunedCells_ || cells < 0.75 * tunedCells_))
    tuneDue_ = true;
  if (!fixed_ && !multiOk && multiIter(shape_.kind)) {
    shape_ = Shape{kTiles, 14};
    tuneDue_ = true;
  }
  auto allowed = [&](int k) {
    return (multiOk || !multiIter(kCand[k].kind)) &&
           (opt_.mode == EP2D_MODE_SPLIT || !wide(kCand[k].kind));
  };
  // a tuning chunk runs a warm-up round and up to kTuneEach timed rounds of
  // every candidate, as many as the chunk holds (at least one timed round)
  // (leaving room, when the chunk is long enough for two rounds besides, for
  // the refinement of the closest candidates below: longer back-to-back runs
  // rank them more reliably than the interleaved rounds)
  constexpr long kTuneEach = 4;
  long roundIters = 0;
  for (int k = 0; k < kNCand; k++)
    if (allowed(k)) roundIters += 2 * ipl(kCand[k].kind);
  const long refineRoom = (long)kRefCands * kRefRounds * kRefLaunches * 4;
  long rounds = std::min<long>(kTuneEach + 1, (c - 2 - refineRoom) / roundIters);
  if (rounds < 2) rounds = std::min<long>(kTuneEach + 1, (c - 2) / roundIters);
  const bool tuning = !fixed_ && tuneDue_ && rounds >= 2;
  ids.push_back(tuning ? 1 : 0);
  ids.push_back(shape_.kind);
  ids.push_back(shape_.rows);
  prepareWork(ids, tuning);
  // RVE table: re-uploaded when an RVE's current buffers changed
  std::vector<int> curs(active_.size());
  for (size_t k = 0; k < active_.size(); k++)
    curs[k] = active_[k].s->cur() | (active_[k].s->vBuffer() << 1);
  if (curs != curs_) {
    for (size_t k = 0; k < active_.size(); k++) {
      hRves_[k] = active_[k].s->dev();
      hRves_[k].cur = active_[k].s->cur();
    }
    EP2D_HIP(hipMemcpyAsync(dRves_, hRves_, active_.size() * sizeof(RveDev),
                            hipMemcpyHostToDevice, stream_));
    curs_ = curs;
  }
  if (nTiles_ == 0) return Seq{0, 0};  // every RVE finishes in the FP64 arithmetic
  long l = 0, d = 0, it = 0;
  if (tuning) {
    // a warm-up round, then the timed rounds, candidates interleaved; a
    // candidate's two launches of a round run back to back between one pair
    // of events (so a multi-iteration candidate leaves V back in its buffer
    // before the next candidate).  After the first timed round the
    // candidates more than a quarter slower than the fastest drop out: the
    // later rounds only rank the close ones (the tuning chunk's iterations
    // on slow shapes cost time)
    bool keep[kNCand];
    for (int k = 0; k < kNCand; k++) keep[k] = allowed(k);
    for (long round = 0; round < rounds; round++) {
      if (round == 2) {
        EP2D_HIP(hipStreamSynchronize(stream_));
        resolveEvents();
        double best = 0.0;
        for (int k = 0; k < kNCand; k++)
          if (keep[k] && tuneN_[k] && (best == 0.0 || st_.candUs[k] < best)) best = st_.candUs[k];
        for (int k = 0; k < kNCand; k++)
          if (keep[k] && tuneN_[k] && st_.candUs[k] > 1.25 * best) keep[k] = false;
      }
      for (int k = 0; k < kNCand; k++) {
        if (!keep[k]) continue;
        const Shape sh = kCand[k];
        const int4* w = dTiles_ + candOff_[k];
        const bool dbl = multiIter(sh.kind);
        Pending ev{};
        if (round > 0) {
          ev = takeEvents();
          ev.cand = k;
          ev.launches = 2 * ipl(sh.kind);   // iterations
          EP2D_HIP(hipEventRecord(ev.a, stream_));
        }
        for (int rep = 0; rep < 2; rep++) {
          launchOne(sh, w, candCnt_[k], dbl, l, d);
          l++;
          if (dbl) d++;
          it += ipl(sh.kind);
        }
        if (round > 0) {
          EP2D_HIP(hipEventRecord(ev.b, stream_));
          evPending_.push_back(ev);
        }
      }
    }
    for (int k = 0; k < kNCand; k++) tuneMs_[k] = 0, tuneN_[k] = 0;
    tunedCells_ = cells;
    tuneDue_ = false;
    st_.tunings++;
    // pick the shape now, so the rest of this chunk already runs it (the
    // tuning launches left d even, as the one-iteration kernels need)
    EP2D_HIP(hipStreamSynchronize(stream_));
    resolveEvents();
    // refine among the candidates within 3% of the best: longer back-to-back
    // runs, so differences of a percent decide reliably
    {
      int near[kNCand], nn = 0;
      const double best = st_.candUs[candIndex(shape_)];
      for (int k = 0; k < kNCand && nn < kRefCands; k++)
        if (allowed(k) && st_.candUs[k] > 0 && st_.candUs[k] <= 1.03 * best) near[nn++] = k;
      long need = 0;
      for (int q = 0; q < nn; q++)
        need += kRefRounds * kRefLaunches * ipl(kCand[near[q]].kind);
      if (nn > 1 && c - it >= need + 2) {
        for (int k = 0; k < kNCand; k++) tuneMs_[k] = 0, tuneN_[k] = 0;
        for (int round = 0; round < kRefRounds; round++)
          for (int q = 0; q < nn; q++) {
            const int k = near[q];
            const Shape sh = kCand[k];
            const bool dbl = multiIter(sh.kind);
            Pending ev = takeEvents();
            ev.cand = k;
            ev.launches = kRefLaunches * ipl(sh.kind);
            EP2D_HIP(hipEventRecord(ev.a, stream_));
            for (int rep = 0; rep < kRefLaunches; rep++) {
              launchOne(sh, dTiles_ + candOff_[k], candCnt_[k], dbl, l, d);
              l++;
              if (dbl) d++;
              it += ipl(sh.kind);
            }
            EP2D_HIP(hipEventRecord(ev.b, stream_));
            evPending_.push_back(ev);
          }
        EP2D_HIP(hipStreamSynchronize(stream_));
        resolveEvents();
      }
    }
    ids.resize(ids.size() - 3);
    ids.push_back(0);
    ids.push_back(shape_.kind);
    ids.push_back(shape_.rows);
    prepareWork(ids, false);
  }
  return runRest(c, it, l, d, bytes, cells, tuning);
}

// Work lists of the admitted RVEs for the current shape (and its single-
// iteration list for a multi-iteration shape, and every candidate's list when
// tuning), uploaded when the RVE set, the tuning state or the shape changed.
void Batch::prepareWork(const std::vector<long long>& ids, bool tuning) {
  if (ids != key_) {
    dropGraphs();
    const Shape aux{kMarch, shape_.rows};
    const Shape esc{kEsc, 8};
    size_t need = (size_t)buildWork(shape_, nullptr) + (size_t)buildWork(esc, nullptr);
    need += (size_t)buildWork(kEscMain, nullptr, true) + (size_t)buildWork(kClose3, nullptr, true) +
            (size_t)buildWork(kClose2, nullptr, true) +
            (size_t)buildWork(Shape{kMarch, kEscMain.rows}, nullptr, true);
    const bool close3 = ipl(shape_.kind) == 4, close2 = ipl(shape_.kind) >= 3;
    if (multiIter(shape_.kind)) need += (size_t)buildWork(aux, nullptr);
    if (close3) need += (size_t)buildWork(kClose3, nullptr);
    if (close2) need += (size_t)buildWork(kClose2, nullptr);
    if (tuning)
      for (int k = 0; k < kNCand; k++) need += (size_t)buildWork(kCand[k], nullptr);
    ensureCapacity(active_.size(), need);
    nTiles_ = buildWork(shape_, hTiles_);
    int off = nTiles_;
    auxOff_ = off;
    auxCnt_ = 0;
    if (multiIter(shape_.kind)) {
      auxCnt_ = buildWork(aux, hTiles_ + off);
      off += auxCnt_;
    }
    rem3Off_ = off;
    rem3Cnt_ = close3 ? buildWork(kClose3, hTiles_ + off) : 0;
    off += rem3Cnt_;
    rem2Off_ = off;
    rem2Cnt_ = close2 ? buildWork(kClose2, hTiles_ + off) : 0;
    off += rem2Cnt_;
    escOff_ = off;
    escCnt_ = buildWork(esc, hTiles_ + off);
    off += escCnt_;
    // the same RVEs on the chain kernels in the FP64 arithmetic, when all of
    // them fit the multi-iteration kernels
    escChains_ = false;
    escMainCnt_ = escC3Cnt_ = escC2Cnt_ = escAuxCnt_ = 0;
    if (escCnt_ > 0) {
      escChains_ = true;
      for (const auto& sl : active_)
        if (sl.s->escalated() && !multi_iter_fits(sl.s->nX(), sl.s->nY())) escChains_ = false;
    }
    if (escChains_) {
      escMainOff_ = off;
      escMainCnt_ = buildWork(kEscMain, hTiles_ + off, true);
      off += escMainCnt_;
      escC3Off_ = off;
      escC3Cnt_ = buildWork(kClose3, hTiles_ + off, true);
      off += escC3Cnt_;
      escC2Off_ = off;
      escC2Cnt_ = buildWork(kClose2, hTiles_ + off, true);
      off += escC2Cnt_;
      escAuxOff_ = off;
      escAuxCnt_ = buildWork(Shape{kMarch, kEscMain.rows}, hTiles_ + off, true);
      off += escAuxCnt_;
    }
    if (tuning)
      for (int k = 0; k < kNCand; k++) {
        candOff_[k] = off;
        candCnt_[k] = buildWork(kCand[k], hTiles_ + off);
        off += candCnt_[k];
      }
    EP2D_HIP(hipMemcpyAsync(dTiles_, hTiles_, (size_t)off * sizeof(int4), hipMemcpyHostToDevice,
                            stream_));
    key_ = ids;
    curs_.clear();
  }
}

// The chunk's iterations after the `it` tuning ones, from launch index l and
// double index d: the main launches (a graph replay when allowed), then for a
// multi-iteration shape its one to four single iterations.
Batch::Seq Batch::runRest(long c, long it, long l, long d, double bytes, double cells,
                          bool tuning) {
  const long rest = c - it;
  const bool timed = opt_.timingEvery > 0;
  const bool two = multiIter(shape_.kind);
  const int n = ipl(shape_.kind);
  long pairs = 0, singles = rest;
  if (two) split_multi(rest, n, pairs, singles);
  const long main = two ? pairs : singles;   // launches of the main kernel
  if (opt_.useGraphs && !tuning) {
    // graph of the main launches; with timing on, the last of them runs
    // outside the graph between dispatch events; a short run (an adaptive
    // check close to the last one) is launched directly, not captured
    long glen = timed && main > 0 ? main - 1 : main;
    if (glen < kMinGraphLaunches) {
      enqueue(glen, two, l, d);
      glen = 0;
    }
    if (glen > 0) {
      const hipGraphExec_t ge = graphFor(glen);
      Pending gv{};
      if (timed) {
        gv = takeEvents();
        gv.graph = true;
        gv.launches = n * glen;
        EP2D_HIP(hipEventRecord(gv.a, stream_));
      }
      EP2D_HIP(hipGraphLaunch(ge, stream_));
      if (timed) {
        EP2D_HIP(hipEventRecord(gv.b, stream_));
        evPending_.push_back(gv);
      }
      l += glen;
      if (two) d += glen;
    }
    if (timed && main > 0) {
      Pending ev = takeEvents();
      ev.bytes = n * bytes;
      ev.launches = n;
      launchOne(shape_, dTiles_, nTiles_, two, l, d, ev.a, ev.b);
      evPending_.push_back(ev);
      l++;
      if (two) d++;
    }
    if (two) closing(singles, l, d);
  } else {
    auto one = [&](bool dbl) {
      const bool t = timed && (st_.launches + l) % opt_.timingEvery == 0;
      Pending ev{};
      if (t) {
        ev = takeEvents();
        ev.bytes = (dbl ? n : 1) * bytes;
        ev.launches = dbl ? n : 1;
      }
      launchOne(shape_, dTiles_, nTiles_, dbl, l, d, ev.a, ev.b);
      l++;
      if (dbl) d++;
      if (t) evPending_.push_back(ev);
    };
    for (long k = 0; k < main; k++) one(two);
    if (two) closing(singles, l, d);   // not timed: only the main kernel is
  }
  st_.launches += l;
  st_.chunks++;
  st_.tilesLaunched += (long long)l * nTiles_;
  st_.cellIterations += (long long)(c * cells);
  st_.bytesAlgorithmic += c * bytes;
  return Seq{l, d};
}

Batch::Pending Batch::takeEvents() {
  if (evPool_.empty()) {
    Pending p{};
    EP2D_HIP(hipEventCreate(&p.a));
    EP2D_HIP(hipEventCreate(&p.b));
    return p;
  }
  Pending p = evPool_.back();
  evPool_.pop_back();
  return p;
}

void Batch::resolveEvents() {
  bool tuned = false;
  for (auto& e : evPending_) {
    float ms = 0.f;
    EP2D_HIP(hipEventElapsedTime(&ms, e.a, e.b));
    if (e.cand >= 0) {
      tuneMs_[e.cand] += ms;
      tuneN_[e.cand] += e.launches;   // iterations
      tuned = true;
    } else if (e.graph) {
      st_.graphMs += ms;
      st_.graphLaunches += e.launches;
    } else {
      st_.timedMs += ms;
      st_.timedBytes += e.bytes;
      st_.timedLaunches += e.launches;
    }
    e.cand = -1;
    e.graph = false;
    evPool_.push_back(e);
  }
  evPending_.clear();
  if (tuned) {
    // the fastest of the candidates timed in this pass (a refinement pass
    // times a few; the others keep their last figures)
    int best = -1;
    double bestUs = 0.0;
    for (int k = 0; k < kNCand; k++) {
      if (!tuneN_[k]) continue;
      const double us = 1e3 * tuneMs_[k] / tuneN_[k];
      st_.candUs[k] = us;
      if (best < 0 || us < bestUs) best = k, bestUs = us;
    }
    if (best >= 0) shape_ = kCand[best];
  }
}

void Batch::step(int* nActive, int* nQueued) {
  EP2D_HIP(hipSetDevice(device_));
  admit();
  if (!active_.empty()) {
    long c = active_.front().s->tsChunk();
    for (auto& sl : active_) c = std::min(c, sl.s->tsChunk());
    c = std::max<long>(c, 1);
    const Seq q = launchChunk(c);
    // the RVEs finishing a time step in the FP64 arithmetic, on their own lists
    const Seq qe = escCnt_ > 0 ? runEsc(c) : Seq{0, 0};
    // the RVEs at a residual check: job at[k] of one reduction launch
    std::vector<int> at(active_.size(), -1);
    int nj = 0;
    for (size_t k = 0; k < active_.size(); k++) {
      Slot& sl = active_[k];
      if (sl.s->escalated()) {
        sl.s->noteBatchIterations(c, qe.launches, qe.doubles);
        st_.cellIterations += (long long)c * sl.s->nX() * sl.s->nY();
        st_.fp64CellIterations += (long long)c * sl.s->nX() * sl.s->nY();
      } else {
        sl.s->noteBatchIterations(c, q.launches, q.doubles);
      }
      if (sl.s->tsAtCheck()) at[k] = nj++;
    }
    if (nj > 0) {
      ensureJobs((size_t)nj);
      const bool f32 = active_.front().s->splitVelocity();
      for (size_t k = 0; k < active_.size(); k++) {
        if (at[k] < 0) continue;
        if (active_[k].s->splitVelocity() != f32)
          throw std::logic_error("batch: velocity storage differs between RVEs");
        hJobs_[at[k]] = active_[k].s->residualJob();
      }
      EP2D_HIP(hipMemcpyAsync(dJobs_, hJobs_, (size_t)nj * sizeof(MaxabsJob),
                              hipMemcpyHostToDevice, stream_));
      launch_maxabs_multi(hJobs_, dJobs_, nj, f32, dRes_, stream_);
      EP2D_HIP(hipMemcpyAsync(hRes_, dRes_, 3 * (size_t)nj * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, stream_));
    }
    EP2D_HIP(hipStreamSynchronize(stream_));
    resolveEvents();
    std::vector<char> gone(active_.size(), 0);
    for (size_t k = 0; k < active_.size(); k++) {
      Slot& sl = active_[k];
      try {
        if (at[k] >= 0) sl.s->tsCheckFinish(hRes_ + 3 * (size_t)at[k]);
        if (sl.s->tsOver()) {
          sl.s->tsFinish();
          if (sl.s->jobDone()) {
            finish(sl);
            gone[k] = 1;
          } else {
            sl.s->tsBegin();
          }
        }
      } catch (const StatusError& e) {
        fail(sl, e.status, e.what());
        gone[k] = 1;
      } catch (const ArgError& e) {
        fail(sl, EP2D_ERR_ARG, e.what());
        gone[k] = 1;
      } catch (const HipError& e) {
        // a failed copy or launch of this RVE's post-processing or next
        // load step fails this request only; slots already finished in this
        // pass are still removed below, so a later step cannot finish them
        // twice
        fail(sl, EP2D_ERR_HIP, e.what());
        gone[k] = 1;
      }
    }
    std::vector<Slot> keep;
    keep.reserve(active_.size());
    for (size_t k = 0; k < active_.size(); k++) {
      if (gone[k]) {
        used_ -= active_[k].bytes;
        active_[k].s.reset();
      } else {
        keep.push_back(std::move(active_[k]));
      }
    }
    active_.swap(keep);
    admit();
  }
  if (nActive) *nActive = (int)active_.size();
  if (nQueued) *nQueued = (int)queue_.size();
}

void Batch::run() {
  int a = 0, q = 0;
  do {
    step(&a, &q);
  } while (a > 0 || q > 0);
}

int Batch::poll(ep2d_batch_result* out, int capacity) {
  const int n = std::min<int>(capacity, (int)finished_.size());
  for (int k = 0; k < n; k++) out[k] = finished_[k];
  finished_.erase(finished_.begin(), finished_.begin() + n);
  return n;
}

void Batch::stats(ep2d_batch_stats* s) const {
  *s = st_;
  s->active = (int)active_.size();
  s->queued = (int)queue_.size();
  s->deviceBytes = used_;
  s->kernel = shape_.kind;
  s->kernelRows = shape_.rows;
  for (const auto& sl : active_) s->fp64Finishes += sl.s->escalations();
  s->nCand = kNCand;
  for (int k = 0; k < kNCand; k++) {
    s->candKind[k] = kCand[k].kind;
    s->candRows[k] = kCand[k].rows;
  }
}

}  // namespace ep2d
