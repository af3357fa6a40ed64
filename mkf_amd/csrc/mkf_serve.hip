// mkf_serve.hip - the serving loop: admission under a device-memory budget,
// host-to-device copies on their own stream, continuous batching of every
// request whose field is on the device into the batched histogram kernel,
// and result delivery. Host code only; the kernels are in mkf_batch.hip.
//
// Threads: callers submit and poll under one mutex; one worker thread owns
// the device work. The worker keeps two launch slots so the next launch is
// queued while the current one runs. Copies of admitted host fields are
// issued by a copier thread on the copy stream: hipMemcpyAsync of a large
// pinned buffer can hold its caller for most of the transfer, which must not
// stall launches and deliveries.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "mkf_arena.hpp"
#include "mkf_internal.h"

using namespace mkf;

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kSlots = 2;
// Host copies queued on the copy stream at once: enough to keep the link
// busy, few enough that waiting host requests hold no device memory.
constexpr int kMaxCopies = 3;

size_t alignUp(size_t v) { return ArenaMap::alignUp(v); }

// The device reservation for copied-in fields; the first-fit bookkeeping is
// ArenaMap's.
class Arena {
 public:
  int init(size_t bytes) {
    map_.reset(bytes);
    if (map_.capacity() == 0) return MKF_OK;
    if (hipMalloc((void**)&base_, map_.capacity()) != hipSuccess) {
      base_ = nullptr;
      map_.reset(0);
      return setError(MKF_ERR_MEM, "serve: cannot reserve %zu device bytes", ArenaMap::alignUp(bytes));
    }
    return MKF_OK;
  }
  void release() {
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
  }
  void* alloc(size_t n) {
    const size_t off = map_.alloc(n);
    return off == ArenaMap::kNone ? nullptr : base_ + off;
  }
  void free(void* p, size_t n) { map_.release((size_t)((char*)p - base_), n); }
  size_t capacity() const { return map_.capacity(); }
  size_t used() const { return map_.used(); }

 private:
  char* base_ = nullptr;
  ArenaMap map_;
};

// A host request whose field does not fit the budget is served as z-slabs
// (sharing one plane, so every plane pair is in exactly one slab); the
// parent sums the slabs' histograms and is reported once, after the last.
struct Parent {
  MkfServeRequest r;         // the caller's request (pBPFD: the caller's output)
  std::vector<size_t> acc;   // nLevels x 256
  int remaining = 0;
  int status = MKF_OK;
  int batchMax = 0;
  double tSubmit = 0, tAdmit = 1e300, tLaunch = 0;
};

struct Request {
  MkfServeRequest r;
  Parent* parent = nullptr;        // set for the slabs of a split request
  std::vector<size_t> slabOut;     // a slab's histograms (r.pBPFD points here)
  BinMapF64 maps[MKF_BATCH_MAX_LEVELS];
  size_t bytes = 0;  // field bytes
  const void* dField = nullptr;
  void* owned = nullptr;  // arena block holding a copied-in field
  hipEvent_t copied = nullptr;    // recorded on the copy stream by the copier
  std::atomic<bool> issued{false};  // the copier has issued the copy (copyStatus is set)
  int copyStatus = MKF_OK;
  hipEvent_t produced = nullptr;  // recorded on the producer's stream at submission
  double tSubmit = 0, tAdmit = 0, tLaunch = 0;
};

struct Slot {
  bool busy = false;
  std::vector<Request*> reqs;
  BatchLauncher* launcher = nullptr;
  unsigned long long* dHist = nullptr;
  unsigned long long* hHist = nullptr;     // mapped pinned
  unsigned long long* hHistDev = nullptr;  // its device address
  hipEvent_t done = nullptr;
  double tLaunch = 0;
  int status = MKF_OK;
};

}  // namespace

struct MkfServer {
  MkfServeConfig cfg{};
  int device = 0;
  Clock::time_point t0;
  Arena arena;
  hipStream_t compute = nullptr, copy = nullptr;
  Slot slots[kSlots];
  std::vector<hipEvent_t> eventPool;

  std::mutex mu;
  std::condition_variable cvWork, cvDone;
  std::deque<Request*> pending;          // guarded by mu
  std::deque<MkfServeResult> results;    // guarded by mu
  MkfServeStats stats{};                 // guarded by mu
  bool stop = false;                     // guarded by mu
  std::deque<Request*> toCopy;           // guarded by mu: admitted host requests for the copier
  bool stopCopier = false;               // guarded by mu
  std::condition_variable cvCopy;
  std::thread worker, copier;

  // Worker-only state.
  std::deque<Request*> copying, ready;
  int admitted = 0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

hipEvent_t takeEvent(MkfServer* S) {
  if (!S->eventPool.empty()) {
    hipEvent_t e = S->eventPool.back();
    S->eventPool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

// Worker thread only (parents are touched by no other thread).
void report(MkfServer* S, Request* q, int status, int batchSize, double tDone) {
  MkfServeResult res;
  res.tag = q->r.tag;
  res.status = status;
  res.batchSize = batchSize;
  res.tSubmit = q->tSubmit;
  res.tAdmit = q->tAdmit;
  res.tLaunch = q->tLaunch;
  res.tDone = tDone;
  Parent* p = q->parent;
  if (p) {
    if (status == MKF_OK) {
      for (size_t i = 0; i < p->acc.size(); ++i) p->acc[i] += q->slabOut[i];
    } else if (p->status == MKF_OK) {
      p->status = status;
    }
    p->tAdmit = std::min(p->tAdmit, q->tAdmit);
    p->tLaunch = std::max(p->tLaunch, q->tLaunch);
    p->batchMax = std::max(p->batchMax, batchSize);
  }
  if (q->owned) S->arena.free(q->owned, q->bytes);
  if (q->copied) S->eventPool.push_back(q->copied);
  if (q->produced) (void)hipEventDestroy(q->produced);
  delete q;
  S->admitted--;
  if (p) {
    if (--p->remaining > 0) return;
    if (p->status == MKF_OK) memcpy(p->r.pBPFD, p->acc.data(), sizeof(size_t) * p->acc.size());
    res.tag = p->r.tag;
    res.status = p->status;
    res.batchSize = p->batchMax;
    res.tSubmit = p->tSubmit;
    res.tAdmit = p->tAdmit;
    res.tLaunch = p->tLaunch;
    delete p;
  }
  std::lock_guard<std::mutex> g(S->mu);
  S->results.push_back(res);
  S->stats.completed++;
}

// Admits queued requests while slots allow. Host requests are admitted in
// arrival order while the copy queue has room and their field fits the
// budget; a device-resident request needs no copy and no budget, so it is
// admitted past host requests that wait (it would otherwise queue behind
// their PCIe transfers).
bool admit(MkfServer* S) {
  std::vector<Request*> fresh;
  {
    std::lock_guard<std::mutex> g(S->mu);
    bool hostBlocked = false;
    int newCopies = 0;
    for (auto it = S->pending.begin(); it != S->pending.end() && S->admitted < S->cfg.maxInFlight;) {
      Request* q = *it;
      if (!(q->r.flags & MKF_SERVE_ON_DEVICE)) {
        if (hostBlocked || (int)S->copying.size() + newCopies >= kMaxCopies) {
          hostBlocked = true;
          ++it;
          continue;
        }
        q->owned = S->arena.alloc(q->bytes);
        if (!q->owned) {  // wait for memory to free up
          hostBlocked = true;
          ++it;
          continue;
        }
        S->stats.heldMax = std::max<uint64_t>(S->stats.heldMax, S->arena.used());
        newCopies++;
      }
      it = S->pending.erase(it);
      S->admitted++;
      fresh.push_back(q);
    }
  }
  const double t = S->now();
  for (Request* q : fresh) {
    q->tAdmit = t;
    if (q->r.flags & MKF_SERVE_ON_DEVICE) {
      q->dField = q->r.pF;
      S->ready.push_back(q);
      continue;
    }
    q->dField = q->owned;
    q->copied = takeEvent(S);
    if (!q->copied) {
      report(S, q, setError(MKF_ERR_HIP, "serve: copy event"), 0, S->now());
      continue;
    }
    {
      std::lock_guard<std::mutex> g(S->mu);
      S->stats.bytesIn += q->bytes;
      S->toCopy.push_back(q);
    }
    S->cvCopy.notify_one();
    S->copying.push_back(q);
  }
  return !fresh.empty();
}

void copierMain(MkfServer* S) {
  (void)hipSetDevice(S->device);
  for (;;) {
    Request* q = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cvCopy.wait(lk, [S] { return S->stopCopier || !S->toCopy.empty(); });
      if (S->toCopy.empty()) return;
      q = S->toCopy.front();
      S->toCopy.pop_front();
    }
    if (hipMemcpyAsync(q->owned, q->r.pF, q->bytes, hipMemcpyHostToDevice, S->copy) != hipSuccess ||
        hipEventRecord(q->copied, S->copy) != hipSuccess)
      q->copyStatus = MKF_ERR_HIP;
    q->issued.store(true, std::memory_order_release);
    S->cvWork.notify_one();
  }
}

// Copies are issued and finish in order on the copy stream.
bool collectCopies(MkfServer* S) {
  bool any = false;
  while (!S->copying.empty()) {
    Request* q = S->copying.front();
    if (!q->issued.load(std::memory_order_acquire)) break;
    if (q->copyStatus == MKF_OK && hipEventQuery(q->copied) == hipErrorNotReady) break;
    S->copying.pop_front();
    any = true;
    if (q->copyStatus != MKF_OK) {
      report(S, q, setError(MKF_ERR_HIP, "serve: field copy"), 0, S->now());
      continue;
    }
    S->ready.push_back(q);
  }
  return any;
}

bool launch(MkfServer* S, Slot& sl) {
  if (S->ready.empty()) return false;
  const int n = std::min<int>(S->cfg.maxBatch, (int)S->ready.size());
  sl.reqs.assign(S->ready.begin(), S->ready.begin() + n);
  S->ready.erase(S->ready.begin(), S->ready.begin() + n);
  std::vector<MkfBatchField> bf(n);
  size_t lv = 0;
  for (int i = 0; i < n; ++i) {
    Request* q = sl.reqs[i];
    MkfBatchField& f = bf[i];
    f.pF = q->dField;
    f.enc = q->r.enc;
    f.def[0] = q->r.def[0];
    f.def[1] = q->r.def[1];
    f.def[2] = q->r.def[2];
    f.nLevels = q->r.nLevels;
    f.pM = q->maps;
    f.pBPFD = (size_t*)(sl.dHist + lv * 256);
    lv += q->r.nLevels;
  }
  sl.tLaunch = S->now();
  for (Request* q : sl.reqs) {
    q->tLaunch = sl.tLaunch;
    if (q->produced) (void)hipStreamWaitEvent(S->compute, q->produced, 0);
  }
  sl.status = batchLauncherRun(sl.launcher, bf.data(), n, &S->cfg.shape, MKF_BATCH_ZERO_OUTPUT, S->compute);
  // The histograms go to the host through a copy kernel: a DMA copy on the
  // compute stream would queue behind field copies on the copy engine.
  if (sl.status == MKF_OK &&
      (copyWords(sl.hHistDev, sl.dHist, lv * 256 * sizeof(unsigned long long), S->compute) != hipSuccess ||
       hipEventRecord(sl.done, S->compute) != hipSuccess))
    sl.status = setError(MKF_ERR_HIP, "serve: result copy");
  sl.busy = true;
  std::lock_guard<std::mutex> g(S->mu);
  S->stats.launches++;
  return true;
}

bool deliver(MkfServer* S, Slot& sl) {
  if (!sl.busy) return false;
  if (sl.status == MKF_OK) {
    const hipError_t e = hipEventQuery(sl.done);
    if (e == hipErrorNotReady) return false;
    if (e != hipSuccess) sl.status = setError(MKF_ERR_HIP, "serve: launch failed");
  }
  const double t = S->now();
  size_t lv = 0;
  const int n = (int)sl.reqs.size();
  for (Request* q : sl.reqs) {
    if (sl.status == MKF_OK)
      memcpy(q->r.pBPFD, sl.hHist + lv * 256, sizeof(size_t) * 256 * (size_t)q->r.nLevels);
    lv += q->r.nLevels;
    report(S, q, sl.status, n, t);
  }
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stats.gpuBusy += t - sl.tLaunch;
  }
  sl.reqs.clear();
  sl.busy = false;
  S->cvDone.notify_all();
  return true;
}

void workerMain(MkfServer* S) {
  (void)hipSetDevice(S->device);
  for (;;) {
    bool moved = admit(S);
    moved |= collectCopies(S);
    for (Slot& sl : S->slots) moved |= deliver(S, sl);
    for (Slot& sl : S->slots)
      if (!sl.busy) moved |= launch(S, sl);
    if (moved) continue;
    const bool gpuWork = !S->copying.empty() || S->slots[0].busy || S->slots[1].busy;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->stop && S->pending.empty() && !gpuWork && S->ready.empty()) break;
    if (gpuWork) {
      // Poll the device again shortly; new submissions wake the worker too.
      S->cvWork.wait_for(lk, std::chrono::microseconds(20));
    } else if (S->pending.empty() || S->admitted >= S->cfg.maxInFlight) {
      S->cvWork.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
}

void freeServer(MkfServer* S) {
  for (Slot& sl : S->slots) {
    if (sl.launcher) batchLauncherDestroy(sl.launcher);
    if (sl.dHist) (void)hipFree(sl.dHist);
    if (sl.hHist) (void)hipHostFree(sl.hHist);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  for (hipEvent_t e : S->eventPool) (void)hipEventDestroy(e);
  if (S->compute) (void)hipStreamDestroy(S->compute);
  if (S->copy) (void)hipStreamDestroy(S->copy);
  S->arena.release();
  delete S;
}

}  // namespace

extern "C" {

MkfServer* mkfServeCreate(const MkfServeConfig* cfg) {
  MkfServer* S = new MkfServer();
  if (cfg) S->cfg = *cfg;
  if (S->cfg.maxInFlight <= 0) S->cfg.maxInFlight = 64;
  if (S->cfg.maxBatch <= 0) S->cfg.maxBatch = 32;
  if (S->cfg.maxBatch > 1024) {
    setError(MKF_ERR_ARG, "serve: maxBatch above 1024");
    delete S;
    return nullptr;
  }
  (void)hipGetDevice(&S->device);
  if (S->cfg.hbmBudget == 0) {
    size_t freeB = 0, totalB = 0;
    (void)hipMemGetInfo(&freeB, &totalB);
    S->cfg.hbmBudget = freeB / 2;
  }
  S->t0 = Clock::now();
  const size_t histBytes = (size_t)S->cfg.maxBatch * MKF_BATCH_MAX_LEVELS * 256 * sizeof(unsigned long long);
  bool ok = S->arena.init(S->cfg.hbmBudget) == MKF_OK &&
            hipStreamCreateWithFlags(&S->compute, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking) == hipSuccess;
  for (Slot& sl : S->slots) {
    if (!ok) break;
    sl.launcher = batchLauncherCreate();
    ok = sl.launcher && hipMalloc((void**)&sl.dHist, histBytes) == hipSuccess &&
         hipHostMalloc((void**)&sl.hHist, histBytes, hipHostMallocMapped) == hipSuccess &&
         hipHostGetDevicePointer((void**)&sl.hHistDev, sl.hHist, 0) == hipSuccess &&
         hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess;
  }
  if (!ok) {
    if (mkfLastError()[0] == 0) setError(MKF_ERR_MEM, "serve: resources");
    freeServer(S);
    return nullptr;
  }
  S->copier = std::thread(copierMain, S);
  S->worker = std::thread(workerMain, S);
  return S;
}

int mkfServeSubmit(MkfServer* S, const MkfServeRequest* r) {
  if (!S || !r) return setError(MKF_ERR_ARG, "serve: no server or request");
  FieldDef d[3];
  const int nd = copyNonZeroDef(d, r->def, 3);
  if (!r->pF || !r->pBPFD || !r->pM || (r->enc != ENC_F32 && r->enc != ENC_F64) || r->nLevels < 1 ||
      r->nLevels > MKF_BATCH_MAX_LEVELS || nd <= 0 || r->def[0] < 0 || r->def[1] < 0 || r->def[2] < 0)
    return setError(MKF_ERR_ARG, "serve: invalid request");
  Request* q = new Request();
  q->r = *r;
  if ((r->flags & MKF_SERVE_ON_DEVICE) && (r->flags & MKF_SERVE_AFTER_STREAM)) {
    if (hipEventCreateWithFlags(&q->produced, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(q->produced, (hipStream_t)r->stream) != hipSuccess) {
      if (q->produced) (void)hipEventDestroy(q->produced);
      delete q;
      return setError(MKF_ERR_HIP, "serve: event on the producer stream");
    }
  }
  memcpy(q->maps, r->pM, sizeof(BinMapF64) * r->nLevels);
  q->r.pM = q->maps;
  const size_t eb = r->enc == ENC_F32 ? 4 : 8;
  q->bytes = (size_t)d[0] * d[1] * d[2] * eb;
  std::vector<Request*> parts{q};
  if (!(r->flags & MKF_SERVE_ON_DEVICE) && alignUp(q->bytes) > S->arena.capacity()) {
    // Split along the slowest non-trivial axis into slabs that share one
    // plane: of at most a quarter of the budget, so several slabs are on the
    // device at once, or, when two planes exceed a quarter, of as many planes
    // as the whole budget holds.
    const int k = nd - 1;
    size_t planeBytes = eb;
    for (int i = 0; i < k; ++i) planeBytes *= (size_t)d[i];
    const long long planes = d[k];
    const long long cap = (long long)S->arena.capacity();
    long long sp = std::min<long long>(planes - 1, cap / 4 / (long long)planeBytes - 1);
    if (sp < 1) sp = std::min<long long>(planes - 1, cap / (long long)planeBytes - 1);
    if (planes < 2 || sp < 1 || alignUp(planeBytes * (size_t)(sp + 1)) > S->arena.capacity()) {
      delete q;
      return setError(MKF_ERR_BUDGET, "serve: field of %zu bytes exceeds the budget of %zu, and so do two planes",
                      alignUp(q->bytes), S->arena.capacity());
    }
    Parent* par = new Parent();
    par->r = *r;
    par->acc.assign((size_t)r->nLevels * 256, 0);
    parts.clear();
    for (long long z0 = 0; z0 < planes - 1; z0 += sp) {
      const long long z1 = std::min<long long>(z0 + sp, planes - 1);
      Request* sq = new Request();
      sq->r = q->r;
      sq->r.pF = (const char*)r->pF + (size_t)z0 * planeBytes;
      for (int i = 0; i < 3; ++i) sq->r.def[i] = d[i];
      sq->r.def[k] = (FieldDef)(z1 - z0 + 1);
      memcpy(sq->maps, q->maps, sizeof(BinMapF64) * r->nLevels);
      sq->r.pM = sq->maps;
      sq->slabOut.assign((size_t)r->nLevels * 256, 0);
      sq->r.pBPFD = sq->slabOut.data();
      sq->bytes = planeBytes * (size_t)(z1 - z0 + 1);
      sq->parent = par;
      parts.push_back(sq);
    }
    par->remaining = (int)parts.size();
    delete q;
  }
  {
    std::lock_guard<std::mutex> g(S->mu);
    if (S->stop) {
      for (Request* x : parts) {
        if (x->produced) (void)hipEventDestroy(x->produced);
        if (x->parent && --x->parent->remaining == 0) delete x->parent;
        delete x;
      }
      return setError(MKF_ERR_ARG, "serve: server is stopping");
    }
    const double t = S->now();
    for (Request* x : parts) {
      x->tSubmit = t;
      if (x->parent) x->parent->tSubmit = t;
      S->pending.push_back(x);
    }
    S->stats.submitted++;
  }
  S->cvWork.notify_one();
  return MKF_OK;
}

int mkfServePoll(MkfServer* S, MkfServeResult* out, int maxOut, int waitMicros) {
  if (!S || !out || maxOut <= 0) return setError(MKF_ERR_ARG, "serve: poll arguments");
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->results.empty() && waitMicros > 0)
    S->cvDone.wait_for(lk, std::chrono::microseconds(waitMicros), [S] { return !S->results.empty(); });
  int n = 0;
  while (n < maxOut && !S->results.empty()) {
    out[n++] = S->results.front();
    S->results.pop_front();
  }
  return n;
}

double mkfServeClock(const MkfServer* S) { return S ? S->now() : 0.0; }

int mkfServeGetStats(MkfServer* S, MkfServeStats* st) {
  if (!S || !st) return setError(MKF_ERR_ARG, "serve: stats arguments");
  std::lock_guard<std::mutex> g(S->mu);
  *st = S->stats;
  return MKF_OK;
}

void mkfServeDestroy(MkfServer* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
  }
  S->cvWork.notify_all();
  if (S->worker.joinable()) S->worker.join();
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stopCopier = true;
  }
  S->cvCopy.notify_all();
  if (S->copier.joinable()) S->copier.join();
  // Results nobody polled are dropped with the server.
  freeServer(S);
}

}  // extern "C"
