// Job engine: many closing jobs on one device at once.
//
// The reference runs one closing at a time from its viewer (GMorpho::Update,
// GMorpho.cu:7775). Here a pool of slots, each one gm_ctx (its own HIP stream
// and buffers) holding one job from loading to copy-back, takes jobs from a
// FIFO queue, driven by a few host worker threads. A job is admitted to a
// free slot when the device memory the slot would hold afterwards fits the
// engine's HBM budget; a job that does not fit yet lets smaller jobs behind
// it go first (backfill). Slots keep their largest buffers, so once warm a
// stream of mixed sizes allocates nothing. Work of different slots overlaps
// on the device. Every job runs exactly the kernels of a single-context run,
// so its outputs are identical to running it alone.
//
// Continuous batching: a worker voxelizes a job on its slot's stream; a
// small grid is then handed to the batcher thread and the worker moves on
// to the next job. The batcher closes and meshes every slot that is ready at
// that moment in one gm_batch_close_mesh call (one kernel launch per step
// for all of them); jobs loaded meanwhile go into the next batch, so batches
// grow with the load. A large grid is closed on its slot's stream,
// concurrently with the batches. The device copies each mesh into pinned
// host memory the caller reads in place or copies out. Batched use submits
// all jobs at once; served use submits them as they arrive.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gmorpho_c.h"
#include "gm_internal.h"

extern "C" uint64_t gm_ctx_device_bytes(const gm_ctx* c);
extern "C" int gm_ctx_reserve(gm_ctx* c, int R);
extern "C" int gm_ctx_release(gm_ctx* c);

namespace {

struct Job {
  gm_job desc{};
  // inputs: copied at submit, or borrowed from the caller (valid until the
  // job is done)
  std::vector<float> P, lattice;
  std::vector<uint32_t> T;
  const float* srcP = nullptr;
  const uint32_t* srcT = nullptr;
  const float* srcLat = nullptr;
  int R = 0;
  int status = GM_OK;
  std::string err;
  bool done = false;
  double tSubmit = 0, tStart = 0, tLoaded = 0, tMeshed = 0, tDone = 0;
  // the mesh: in a pinned block of the result pool (DMA target, no host
  // copy) or, for very large meshes, in pageable vectors
  void* blk = nullptr;
  int blkCls = -1;
  std::vector<float> bigV, bigN;
  std::vector<uint32_t> bigT;
  float *V = nullptr, *Nrm = nullptr;
  uint32_t* Tri = nullptr;
  uint32_t nv = 0, nt = 0;
  int slot = -1;
};

// Pinned host blocks for finished meshes, in power-of-two size classes from
// 64 KiB to 256 MiB, reused after a job is released: the device copies the
// mesh straight into the block a caller then reads (gm_engine_view) or
// copies out once (gm_engine_take). hipHostMalloc is slow and pins pages, so
// blocks are kept until the engine is destroyed; larger meshes go to
// pageable memory (a pageable copy runs at a fraction of the DMA rate).
struct ResultPool {
  static constexpr int kClasses = 13;
  static constexpr size_t kMin = size_t(64) << 10;
  std::vector<void*> free[kClasses];
  std::vector<std::pair<void*, int>> all;
  std::mutex m;
  static int cls_of(size_t bytes) {
    size_t sz = kMin;
    for (int k = 0; k < kClasses; ++k, sz <<= 1)
      if (bytes <= sz) return k;
    return -1;
  }
  void* get(int k) {
    {
      std::lock_guard<std::mutex> lk(m);
      if (!free[k].empty()) {
        void* p = free[k].back();
        free[k].pop_back();
        return p;
      }
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, kMin << k, hipHostMallocDefault) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(m);
    all.push_back({p, k});
    return p;
  }
  void put(void* p, int k) {
    std::lock_guard<std::mutex> lk(m);
    free[k].push_back(p);
  }
  void clear() {
    for (auto& a : all) hipHostFree(a.first);
    all.clear();
    for (auto& f : free) f.clear();
  }
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// One job in flight: a context (its own HIP stream and buffers) and the job
// it holds from loading to copy-back.
struct Slot {
  gm_ctx* ctx = nullptr;
  int capR = 0;            // largest base resolution this context has run
  uint64_t reserved = 0;   // device bytes the context holds (estimate, then measured)
  bool busy = false;
  Job* job = nullptr;
  int se = 0;              // structuring element of the loaded job
  int closeRc = 0;         // outcome of its batched closing
  std::string closeErr;
};

// Host threads that load jobs and copy meshes back; the contexts (slots) can
// be many more, so that jobs keep loading while earlier ones wait for a batch.
constexpr int kMaxWorkers = 16;

// Jobs up to this base resolution are closed in batches; larger grids keep
// the GPU busy on their own.
constexpr int kBatchMaxRes = 192;

// Device bytes of one context whose buffers are sized for base resolution R
// (the buffers alloc_grid, the mesh output and the radius field allocate),
// running job j: plus its mesh input and, for a radius above the chord
// tables, the distance-transform scratch of its own grid.
uint64_t context_bytes(int R, int L, const gm_job& j) {
  const uint64_t N = 2ull * (uint64_t)R, W = N / 32;
  uint64_t P = 1;
  while (P < (uint64_t)R) P <<= 1;
  uint64_t b = 0;
  b += 6 * N * N * W * 4;                         // I, D, K, C, voxelizer marks and shell
  b += 6 * (uint64_t)R * R * W * 2;               // DF, TG, CL, UG, I-full and D-any cells
  b += 2 * (N / 8) * (N / 8) * W;                 // brick summaries
  b += (uint64_t)R * R * R;                       // cell-byte staging
  b += 7 * (P / 8) * (P / 8) * (P / 8) * 4;       // MC block counts, list positions, emit order
  const uint64_t listCap = std::min<uint64_t>((P / 8) * (P / 8) * (P / 8), std::max<uint64_t>(64, 8 * (P / 8) * (P / 8)));
  b += listCap * 512 * 12;                        // MC cube cases and cell bases of listed blocks (initial)
  b += 4 * N * N * 24 + 8 * N * N * 12;           // mesh output (initial capacity)
  b += gm::tile_queue_words(R) * 4;               // active-tile queues
  if (L > 0) {
    b += (uint64_t)L * L * L * 4;                         // radius lattice
    b += (gm::radius_plane_words(R) + (uint64_t)R * R * R / 4) * 4;  // radius planes, masked K
    b += 2 * (gm::brick_words(R) + 1) * 4;                // value sets, K brick list
  }
  if (j.radius > 31 && j.radius <= 127) {
    const uint64_t Nj = 2ull * (uint64_t)j.base_res;
    b += 8 * Nj * Nj * Nj;                        // distance-transform scratch (radii above the chord tables)
  }
  b += (uint64_t)j.num_of_vertices * 24 + (uint64_t)j.num_of_faces * 12;
  return b + (16ull << 20);                       // tables, scans, slack
}

}  // namespace

struct gm_engine {
  int device = 0;
  uint64_t budget = 0;
  std::vector<Slot> slots;
  std::vector<std::thread> workers;
  std::mutex m;
  std::condition_variable cvWork, cvDone, cvReady;
  std::deque<Job*> queue;
  std::unordered_map<uint64_t, Job*> jobs;
  uint64_t nextId = 1;
  uint64_t reservedTotal = 0;
  int running = 0;
  bool stop = false;
  std::string err;
  // Continuous batching: a worker loads a small job into a free slot and
  // hands the slot to the batcher (`ready`), then goes on to the next job;
  // the batcher closes and meshes everything ready in one
  // gm_batch_close_mesh call (one launch per step for all of them) and puts
  // the slots in `closed`, where the workers copy the meshes back. Jobs
  // loaded while a batch runs join the next one. All under m.
  gm_batch* batch = nullptr;
  std::thread batcher;
  std::vector<Slot*> ready;
  std::deque<Slot*> closed;
  int batchedPending = 0;  // slots handed to the batcher, not yet in `closed`
  int evicting = 0;        // idle slots whose buffers a worker is releasing
  bool bstop = false;
  ResultPool pool;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // Buffer size of a context running j: a context that batches is sized
  // once for the largest batched grid the budget allows every slot to hold
  // (reserveRes, set at creation), so jobs of mixed small sizes never
  // reallocate.
  int reserveRes = kBatchMaxRes;
  int buffer_res(const Job* j) const { return std::max(j->R, j->R <= kBatchMaxRes ? reserveRes : j->R); }

  // First queued job that some free slot can run within the budget, on the
  // free slot whose reservation grows least (caller holds m). A job that
  // does not fit yet lets smaller jobs behind it go first.
  bool pick(Slot*& out, Job*& job) {
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      Job* j = *it;
      Slot* best = nullptr;
      uint64_t bestGrow = 0, bestNeed = 0;
      const int Rj = buffer_res(j);
      for (Slot& s : slots) {
        if (s.busy) continue;
        const int R = std::max(s.capR, Rj);
        const uint64_t need = std::max(s.reserved, context_bytes(R, j->desc.lattice_res, j->desc));
        const uint64_t grow = need - s.reserved;
        if (!best || grow < bestGrow || (grow == bestGrow && s.capR < best->capR)) {
          best = &s;
          bestGrow = grow;
          bestNeed = need;
        }
      }
      if (!best) return false;  // no free slot
      if (reservedTotal + bestGrow > budget) continue;
      reservedTotal += bestGrow;
      best->reserved = bestNeed;
      best->capR = std::max(best->capR, Rj);
      best->busy = true;
      best->job = j;
      queue.erase(it);
      out = best;
      job = j;
      return true;
    }
    return false;
  }

  bool step(Slot& s, Job* j, int rc) {
    if (rc != GM_OK && j->status == GM_OK) {
      j->status = rc;
      j->err = gm_last_error(s.ctx);
    }
    return rc == GM_OK;
  }

  // Voxelize the job and set its radius field; a small grid then goes to
  // the batcher (true), a large one is closed and meshed on the slot's own
  // stream, concurrently with the batches (false). Errors end the job (false).
  bool load(Slot& s, Job* j, bool& batched) {
    gm_ctx* c = s.ctx;
    const gm_job& d = j->desc;
    batched = false;
    if (!j->srcP) {
      j->srcP = j->P.data();
      j->srcT = j->T.data();
      j->srcLat = j->lattice.empty() ? nullptr : j->lattice.data();
    }
    if (!step(s, j, gm_ctx_reserve(c, s.capR))) return false;
    // gm_load uploads through the context's own pinned staging
    if (!step(s, j, gm_load(c, j->srcP, nullptr, d.num_of_vertices, j->srcT, d.num_of_faces, d.base_res, d.se_size)))
      return false;
    if (j->srcLat) {
      if (!step(s, j, gm_scale_lattice_f32(c, j->srcLat, d.lattice_res))) return false;
    } else if (!step(s, j, gm_scale_uniform(c, d.radius))) {
      return false;
    }
    j->tLoaded = now();
    if (d.base_res <= kBatchMaxRes) {
      s.se = d.se;
      batched = true;
      return true;
    }
    return step(s, j, gm_update(c, d.se));
  }

  // Mesh counts, then the mesh copied by the device into a pinned result
  // block (or pageable vectors for very large meshes).
  void finish(Slot& s, Job* j) {
    gm_ctx* c = s.ctx;
    if (!step(s, j, gm_mesh_counts(c, &j->nv, &j->nt))) return;
    j->tMeshed = now();
    const size_t bv = align256((size_t)j->nv * 12), bt = align256((size_t)j->nt * 12);
    const int k = ResultPool::cls_of(2 * bv + bt);
    if (k >= 0) {
      j->blk = pool.get(k);
      if (!j->blk) {
        j->status = GM_ERR_HIP;
        j->err = "GMorpho Error: pinned host buffer allocation failed";
        return;
      }
      j->blkCls = k;
      j->V = (float*)j->blk;
      j->Nrm = (float*)((char*)j->blk + bv);
      j->Tri = (uint32_t*)((char*)j->blk + 2 * bv);
    } else {
      j->bigV.resize(3 * (size_t)j->nv);
      j->bigN.resize(3 * (size_t)j->nv);
      j->bigT.resize(3 * (size_t)j->nt);
      j->V = j->bigV.data();
      j->Nrm = j->bigN.data();
      j->Tri = j->bigT.data();
    }
    step(s, j, gm_mesh_copy(c, j->V, j->Nrm, j->Tri));
  }

  // The job is done and its slot free (caller holds m). The slot's context
  // may have grown past its estimate (mesh output, block lists): it is
  // charged what it holds.
  void complete(Slot& s, Job* j) {
    const uint64_t held = gm_ctx_device_bytes(s.ctx);
    if (held > s.reserved) {
      reservedTotal += held - s.reserved;
      s.reserved = held;
    }
    s.busy = false;
    s.job = nullptr;
    --running;
    // a borrowed input is the caller's again
    j->srcP = nullptr;
    j->srcT = nullptr;
    j->srcLat = nullptr;
    j->tDone = now();
    j->done = true;
    cvDone.notify_all();
    cvWork.notify_all();
  }

  // Give a finished job's result block back (caller holds m).
  void release_job(Job* j) {
    if (j->blk) pool.put(j->blk, j->blkCls);
    delete j;
  }

  void batch_loop() {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      // everything ready goes at once (holding a batch back until more jobs
      // are loaded left the device idle and measured slower)
      cvReady.wait(lk, [&] { return bstop || !ready.empty(); });
      if (ready.empty()) return;  // stopping
      std::vector<Slot*> take;
      take.swap(ready);
      lk.unlock();
      // one call per structuring element present
      for (int se : {GM_SE_SPHERE, GM_SE_CUBE}) {
        std::vector<gm_ctx*> ctxs;
        std::vector<Slot*> group;
        for (Slot* s : take)
          if (s->se == se) {
            ctxs.push_back(s->ctx);
            group.push_back(s);
          }
        if (group.empty()) continue;
        const int rc = gm_batch_close_mesh(batch, ctxs.data(), (int)ctxs.size(), se);
        for (Slot* s : group) {
          s->closeRc = rc;
          s->closeErr.clear();
        }
        if (rc != GM_OK) {
          // a job of the batch failed: each job's own status comes from
          // closing it alone (the loaded grids are unchanged by a closing)
          for (Slot* s : group) {
            s->closeRc = gm_update(s->ctx, se);
            if (s->closeRc != GM_OK) s->closeErr = gm_last_error(s->ctx);
          }
        }
      }
      lk.lock();
      for (Slot* s : take) closed.push_back(s);
      batchedPending -= (int)take.size();
      cvWork.notify_all();
    }
  }

  void worker() {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      Slot* s = nullptr;
      Job* j = nullptr;
      Slot* evict = nullptr;
      // closed batches first: copying their meshes back frees their slots
      cvWork.wait(lk, [&] {
        if (!closed.empty()) return true;
        if (stop) return batchedPending == 0;
        for (;;) {
          if (pick(s, j)) return true;
          // jobs wait, none fits, and no slot holds a job: the buffers the
          // idle slots keep are what blocks the budget; release the largest
          if (queue.empty() || running != 0) return false;
          for (Slot& t : slots)
            if (!t.busy && t.capR > 0 && (!evict || t.reserved > evict->reserved)) evict = &t;
          if (evict) {
            evict->busy = true;
            ++evicting;
            return true;
          }
          // another worker is still releasing a slot: wait for it
          if (evicting) return false;
          // nothing left to release: the first queued job cannot run here
          Job* jf = queue.front();
          queue.pop_front();
          jf->status = GM_ERR_CAPACITY;
          jf->err = "GMorpho Error: job does not fit the engine's HBM budget beside the buffers other slots keep";
          jf->srcP = nullptr;
          jf->srcT = nullptr;
          jf->srcLat = nullptr;
          jf->tStart = jf->tLoaded = jf->tMeshed = jf->tDone = now();
          jf->done = true;
          cvDone.notify_all();
        }
      });
      if (evict) {
        lk.unlock();
        gm_ctx_release(evict->ctx);
        const uint64_t held = gm_ctx_device_bytes(evict->ctx);
        lk.lock();
        reservedTotal -= evict->reserved;
        evict->reserved = held;
        reservedTotal += held;
        evict->capR = 0;
        evict->busy = false;
        --evicting;
        cvWork.notify_all();
        continue;
      }
      if (!closed.empty()) {
        s = closed.front();
        closed.pop_front();
        j = s->job;
        lk.unlock();
        if (s->closeRc != GM_OK) {
          j->status = s->closeRc;
          j->err = s->closeErr;
        } else {
          finish(*s, j);
        }
        lk.lock();
        complete(*s, j);
        continue;
      }
      if (!s) return;  // stopping, nothing left in the batcher
      j->slot = (int)(s - slots.data());
      j->tStart = now();
      ++running;
      lk.unlock();
      bool batched = false;
      const bool ok = load(*s, j, batched);
      lk.lock();
      if (ok && batched) {
        ready.push_back(s);
        ++batchedPending;
        cvReady.notify_one();
        continue;
      }
      if (ok) {
        lk.unlock();
        finish(*s, j);
        lk.lock();
      }
      complete(*s, j);
    }
  }
};

namespace {
int efail(gm_engine* e, int code, const std::string& msg) {
  e->err = "GMorpho Error: " + msg;
  return code;
}
}  // namespace

extern "C" {

int gm_engine_create(int device, int slots, uint64_t hbm_budget_bytes, gm_engine** out) {
  if (!out || slots < 1 || slots > 256) return GM_ERR_ARG;
  *out = nullptr;
  gm_engine* e = new gm_engine();
  e->device = device;
  e->budget = hbm_budget_bytes;
  e->slots.resize(slots);
  {
    gm_job none{};
    while (e->reserveRes > 16 && (uint64_t)slots * context_bytes(e->reserveRes, 0, none) > hbm_budget_bytes)
      e->reserveRes -= 16;
  }
  for (auto& s : e->slots) {
    if (gm_create(device, &s.ctx) == GM_OK) {
      // a slot runs a different job almost every time: capture only what repeats
      gm_set_graphs(s.ctx, 2);
    } else {
      for (auto& t : e->slots)
        if (t.ctx) gm_destroy(t.ctx);
      delete e;
      return GM_ERR_HIP;
    }
  }
  if (gm_batch_create(device, &e->batch) != GM_OK) {
    for (auto& t : e->slots)
      if (t.ctx) gm_destroy(t.ctx);
    delete e;
    return GM_ERR_HIP;
  }
  gm_batch_set_graphs(e->batch, 2);
  e->batcher = std::thread([e] { e->batch_loop(); });
  const int nw = std::min(slots, kMaxWorkers);
  for (int i = 0; i < nw; ++i) e->workers.emplace_back([e] { e->worker(); });
  *out = e;
  return GM_OK;
}

int gm_engine_destroy(gm_engine* e) {
  if (!e) return GM_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(e->m);
    e->stop = true;
    // queued jobs are dropped; running ones finish first
    e->queue.clear();
  }
  e->cvWork.notify_all();
  // the workers finish the jobs they hold, and those in the batcher
  for (auto& t : e->workers)
    if (t.joinable()) t.join();
  {
    std::lock_guard<std::mutex> lk(e->m);
    e->bstop = true;
  }
  e->cvReady.notify_all();
  if (e->batcher.joinable()) e->batcher.join();
  if (e->batch) gm_batch_destroy(e->batch);
  for (auto& s : e->slots)
    if (s.ctx) gm_destroy(s.ctx);
  for (auto& kv : e->jobs) delete kv.second;
  e->pool.clear();
  delete e;
  return GM_OK;
}

const char* gm_engine_last_error(const gm_engine* e) { return e ? e->err.c_str() : ""; }

}  // extern "C"

namespace {
int submit_job(gm_engine* e, const gm_job* job, uint64_t* id, bool borrow) {
  if (!e || !job || !id) return GM_ERR_ARG;
  const gm_job& d = *job;
  if (!d.P || !d.T || d.num_of_vertices <= 0 || d.num_of_faces <= 0) return efail(e, GM_ERR_ARG, "empty mesh");
  if (d.base_res < 16 || d.base_res > 1024 || (d.base_res % 16) != 0)
    return efail(e, GM_ERR_UNSUPPORTED, "base resolution must be a multiple of 16 in [16,1024]");
  if (d.lattice && d.lattice_res < 2) return efail(e, GM_ERR_ARG, "lattice resolution below 2");
  Job* j = new Job();
  j->desc = d;
  j->R = d.base_res;
  if (!d.lattice) j->desc.lattice_res = 0;
  if (borrow) {
    j->srcP = d.P;
    j->srcT = d.T;
    j->srcLat = d.lattice;
  } else {
    j->P.assign(d.P, d.P + 3 * (size_t)d.num_of_vertices);
    j->T.assign(d.T, d.T + 3 * (size_t)d.num_of_faces);
    if (d.lattice) j->lattice.assign(d.lattice, d.lattice + (size_t)d.lattice_res * d.lattice_res * d.lattice_res);
  }
  j->desc.P = nullptr;
  j->desc.T = nullptr;
  j->desc.lattice = nullptr;
  // what the job needs on a fresh slot (the sizing pick uses): a job that
  // could never be admitted is refused here rather than left queued
  if (context_bytes(e->buffer_res(j), j->desc.lattice_res, j->desc) > e->budget) {
    delete j;
    return efail(e, GM_ERR_CAPACITY, "job does not fit the engine's HBM budget");
  }
  {
    std::lock_guard<std::mutex> lk(e->m);
    if (e->stop) {
      delete j;
      return efail(e, GM_ERR_STATE, "engine is stopping");
    }
    j->tSubmit = e->now();
    *id = e->nextId++;
    e->jobs[*id] = j;
    e->queue.push_back(j);
  }
  e->cvWork.notify_all();
  return GM_OK;
}
}  // namespace

extern "C" {

int gm_engine_submit(gm_engine* e, const gm_job* job, uint64_t* id) { return submit_job(e, job, id, false); }

int gm_engine_submit_borrowed(gm_engine* e, const gm_job* job, uint64_t* id) { return submit_job(e, job, id, true); }

int gm_engine_wait(gm_engine* e, uint64_t id, uint32_t* nv, uint32_t* nt) {
  if (!e) return GM_ERR_ARG;
  std::unique_lock<std::mutex> lk(e->m);
  auto it = e->jobs.find(id);
  if (it == e->jobs.end()) return efail(e, GM_ERR_ARG, "unknown job id");
  Job* j = it->second;
  e->cvDone.wait(lk, [&] { return j->done; });
  if (j->status != GM_OK) {
    e->err = j->err;
    return j->status;
  }
  if (nv) *nv = j->nv;
  if (nt) *nt = j->nt;
  return GM_OK;
}

int gm_engine_take(gm_engine* e, uint64_t id, float* V, float* Nrm, uint32_t* T) {
  if (!e) return GM_ERR_ARG;
  Job* j = nullptr;
  {
    std::unique_lock<std::mutex> lk(e->m);
    auto it = e->jobs.find(id);
    if (it == e->jobs.end()) return efail(e, GM_ERR_ARG, "unknown job id");
    j = it->second;
    e->cvDone.wait(lk, [&] { return j->done; });
    e->jobs.erase(it);
  }
  int rc = j->status;
  if (rc == GM_OK) {
    if (V) std::memcpy(V, j->V, (size_t)j->nv * 12);
    if (Nrm) std::memcpy(Nrm, j->Nrm, (size_t)j->nv * 12);
    if (T) std::memcpy(T, j->Tri, (size_t)j->nt * 12);
  } else {
    e->err = j->err;
  }
  std::lock_guard<std::mutex> lk(e->m);
  e->release_job(j);
  return rc;
}

int gm_engine_view(gm_engine* e, uint64_t id, const float** V, const float** Nrm, const uint32_t** T, uint32_t* nv,
                   uint32_t* nt) {
  if (!e) return GM_ERR_ARG;
  std::unique_lock<std::mutex> lk(e->m);
  auto it = e->jobs.find(id);
  if (it == e->jobs.end()) return efail(e, GM_ERR_ARG, "unknown job id");
  Job* j = it->second;
  e->cvDone.wait(lk, [&] { return j->done; });
  if (j->status != GM_OK) {
    e->err = j->err;
    return j->status;
  }
  if (V) *V = j->V;
  if (Nrm) *Nrm = j->Nrm;
  if (T) *T = j->Tri;
  if (nv) *nv = j->nv;
  if (nt) *nt = j->nt;
  return GM_OK;
}

int gm_engine_release(gm_engine* e, uint64_t id) {
  if (!e) return GM_ERR_ARG;
  std::unique_lock<std::mutex> lk(e->m);
  auto it = e->jobs.find(id);
  if (it == e->jobs.end()) return efail(e, GM_ERR_ARG, "unknown job id");
  Job* j = it->second;
  e->cvDone.wait(lk, [&] { return j->done; });
  e->jobs.erase(it);
  e->release_job(j);
  return GM_OK;
}

int gm_engine_job_phases(gm_engine* e, uint64_t id, double t[5]) {
  if (!e || !t) return GM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->m);
  auto it = e->jobs.find(id);
  if (it == e->jobs.end()) return efail(e, GM_ERR_ARG, "unknown job id");
  const Job* j = it->second;
  t[0] = j->tSubmit;
  t[1] = j->tStart;
  t[2] = j->tLoaded;
  t[3] = j->tMeshed;
  t[4] = j->tDone;
  return GM_OK;
}

int gm_engine_job_times(gm_engine* e, uint64_t id, double t[3]) {
  if (!e || !t) return GM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->m);
  auto it = e->jobs.find(id);
  if (it == e->jobs.end()) return efail(e, GM_ERR_ARG, "unknown job id");
  t[0] = it->second->tSubmit;
  t[1] = it->second->tStart;
  t[2] = it->second->tDone;
  return GM_OK;
}

int gm_engine_now(gm_engine* e, double* t) {
  if (!e || !t) return GM_ERR_ARG;
  *t = e->now();
  return GM_OK;
}

int gm_engine_stats(gm_engine* e, uint64_t* reserved_bytes, int* running, int* queued) {
  if (!e) return GM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->m);
  if (reserved_bytes) *reserved_bytes = e->reservedTotal;
  if (running) *running = e->running;
  if (queued) *queued = (int)e->queue.size();
  return GM_OK;
}

}  // extern "C"
