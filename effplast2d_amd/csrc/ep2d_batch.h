// ep2d_batch.h -- many RVEs on one GPU: one launch per iteration over the
// tiles of every admitted RVE, admission by device-memory budget, retirement
// on convergence and backfill from a queue.
#pragma once

#include <chrono>
#include <deque>
#include <memory>
#include <vector>

#include "ep2d_solver.h"

namespace ep2d {

class Batch {
 public:
  explicit Batch(const ep2d_batch_options& opt);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  long long submit(const ep2d_request& rq);
  void step(int* nActive, int* nQueued);
  void run();
  int poll(ep2d_batch_result* out, int capacity);
  double clock() const;
  void stats(ep2d_batch_stats* s) const;

 private:
  struct Job {
    long long id;
    ep2d_request rq;
    double tSubmit;
  };
  struct Slot {
    std::unique_ptr<Solver> s;
    long long id = 0, tag = 0;
    double tSubmit = 0, tAdmit = 0;
    size_t bytes = 0;
  };
  struct Pending {
    hipEvent_t a = nullptr, b = nullptr;
    double bytes = 0;  // algorithmic bytes of the timed launches
    long launches = 0;   // iterations covered by the timed launch(es)
    int cand = -1;     // shape candidate under tuning, -1 for statistics
    bool graph = false;  // a whole graph replay (launches counted in graphLaunches)
  };
  static constexpr int kTiles = 1, kMarch = 2, kPersist = 3, kMarch2 = 4, kPair = 5,
                       kTriple = 6, kQuad = 7, kWide3 = 8, kWide4 = 9;
  // the chains with two waves side by side per stage
  static constexpr int kTriple2 = 11, kQuad2 = 12;
  // the list of the RVEs finishing a time step in the FP64 arithmetic (tiles
  // of 8 rows for the batched tile kernel); not a shape the tuner times
  static constexpr int kEsc = 10;
  // iterations per launch of a kernel kind; a kind with more than one runs
  // the chunk's last iterations through the one-iteration marching kernel
  static int ipl(int kind) {
    return kind == kQuad || kind == kWide4 || kind == kQuad2 ? 4
           : kind == kTriple || kind == kWide3 || kind == kTriple2 ? 3
           : (kind == kMarch2 || kind == kPair ? 2 : 1);
  }
  static bool chain(int kind) {
    return kind == kTriple || kind == kQuad || kind == kTriple2 || kind == kQuad2;
  }
  // waves per stage of a chain kind
  static int stageWaves(int kind) { return kind == kTriple2 || kind == kQuad2 ? 2 : 1; }
  static bool wide(int kind) { return kind == kWide3 || kind == kWide4; }
  static bool multiIter(int kind) { return ipl(kind) > 1; }
  struct Shape {
    int kind = kTiles;  // fused tile, row-marching, persistent tile, two-iteration
                        // marching on one wave or on a pair of waves,
                        // three- or four-iteration marching on a chain of
                        // three or four stages, each on one or two waves, or
                        // on one wave with two columns per lane
    int rows = 8;       // tile height or segment rows
  };
  static constexpr int kNCand = 11;
  // refinement of the tuning: up to kRefCands candidates within 3% of the
  // best, kRefRounds rounds of kRefLaunches back-to-back launches each
  static constexpr int kRefCands = 3, kRefRounds = 4, kRefLaunches = 4;
  static const Shape kCand[kNCand];
  static const Shape kClose3, kClose2;  // launches that close a chunk (closing())
  static const Shape kEscMain;          // chains of the FP64 finish (runEsc())
  static_assert(kNCand <= EP2D_MAX_CAND, "candidate table exceeds the stats record");
  struct GraphEntry {
    long len = 0;
    long lastUse = 0;   // chunk counter at the last replay (least recently used goes)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  static constexpr int kMaxGraphs = 6;
  static constexpr long kMinGraphLaunches = 16;  // shorter runs launch directly

  size_t estimateBytes(const ep2d_request& rq) const;
  void admit();
  void fail(Slot& sl, int status, const char* what);
  void finish(Slot& sl);
  struct Seq {
    long launches, doubles;
  };
  Seq launchChunk(long c);
  void prepareWork(const std::vector<long long>& ids, bool tuning);
  static int candIndex(Shape sh);
  Seq runRest(long c, long it, long l, long d, double bytes, double cells, bool tuning);
  // work list of the admitted RVEs for a shape: those advancing in their own
  // scheme, or (esc) those finishing a time step in the FP64 arithmetic
  int buildWork(Shape sh, int4* out, bool esc = false) const;
  Seq runEsc(long c);  // c iterations of the RVEs in the FP64 arithmetic
  void launchOne(Shape sh, const int4* work, int n, bool dbl, long l, long d,
                 hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
  void enqueue(long n, bool dbl, long& l, long& d);
  void closing(long singles, long& l, long& d);
  void ensureCapacity(size_t nRves, size_t nTiles);
  void dropGraphs();
  hipGraphExec_t graphFor(long len);
  Pending takeEvents();
  void resolveEvents();

  ep2d_batch_options opt_;
  int device_ = 0;
  MemPool pool_;   // buffers of retired RVEs, reused by the next admissions
  hipStream_t stream_ = nullptr;
  std::chrono::steady_clock::time_point t0_;
  std::deque<Job> queue_;
  std::vector<Slot> active_;
  std::vector<ep2d_batch_result> finished_;
  long long nextId_ = 1;
  size_t budget_ = 0, used_ = 0;

  // device tables of the batched launch (pinned host staging); while the
  // kernel shape is being tuned the work lists of all candidates follow the
  // current one, candidate k at candOff_[k]
  RveDev* dRves_ = nullptr;
  int4* dTiles_ = nullptr;
  RveDev* hRves_ = nullptr;
  // the residual checks of one chunk, reduced in one launch: job table
  // (pinned host and device copies) and three result words per job
  MaxabsJob* hJobs_ = nullptr;
  MaxabsJob* dJobs_ = nullptr;
  unsigned long long* hRes_ = nullptr;
  unsigned long long* dRes_ = nullptr;
  size_t capJobs_ = 0;
  void ensureJobs(size_t n);
  int4* hTiles_ = nullptr;
  size_t capRves_ = 0, capTiles_ = 0;
  int nTiles_ = 0;
  Shape shape_{};
  bool fixed_ = false;
  std::vector<long long> key_;    // RVE ids, tuning flag and shape of the work list
  std::vector<int> curs_;         // current buffers in the uploaded RVE table
  int candOff_[kNCand] = {}, candCnt_[kNCand] = {};
  int auxOff_ = 0, auxCnt_ = 0;   // single-iteration segments of a two-iteration shape
  // segments of the three- and two-iteration launches that close a chunk of
  // a four- or three-iteration shape (closing())
  int rem3Off_ = 0, rem3Cnt_ = 0, rem2Off_ = 0, rem2Cnt_ = 0;
  int escOff_ = 0, escCnt_ = 0;   // tiles of the RVEs in the FP64 arithmetic
  // their chain segments (four stages on two waves), closing three- and two-
  // iteration segments and single-iteration segments, when every one of
  // them fits the multi-iteration kernels (escChains_)
  int escMainOff_ = 0, escMainCnt_ = 0, escC3Off_ = 0, escC3Cnt_ = 0;
  int escC2Off_ = 0, escC2Cnt_ = 0, escAuxOff_ = 0, escAuxCnt_ = 0;
  bool escChains_ = false;

  // kernel-shape tuning on real iterations
  bool tuneDue_ = true;
  double tunedCells_ = 0;
  double tuneMs_[kNCand] = {};
  long tuneN_[kNCand] = {};

  // graphs of chunk lengths seen since the tables were last uploaded
  std::vector<GraphEntry> graphs_;

  // statistics
  ep2d_batch_stats st_{};
  std::vector<Pending> evPool_;
  std::vector<Pending> evPending_;
};

}  // namespace ep2d
