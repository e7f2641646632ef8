// Job engine: many closing jobs on one device at once.
//
// The reference runs one closing at a time from its viewer (GMorpho::Update,
// GMorpho.cu:7775). Here a pool of slots, each one gm_ctx (its own HIP stream
// and buffers) driven by its own host thread, takes jobs from a FIFO queue.
// A job is admitted to a free slot when the device memory the slot would hold
// afterwards fits the engine's HBM budget; a job that does not fit yet lets
// smaller jobs behind it go first (backfill). Slots keep their largest
// buffers, so once warm a stream of mixed sizes allocates nothing. Work of
// different slots overlaps on the device. Every job runs exactly the kernels
// of a single-context run, so its outputs are identical to running it alone.
//
// Continuous batching: a slot voxelizes its job on its own stream; a small
// grid is then handed to the batcher thread, which closes and meshes every
// context that is ready at that moment in one gm_batch_close_mesh call (one
// kernel launch per step for all of them), while jobs that arrive meanwhile
// are loaded by the other slots and go into the next batch. A large grid is
// closed on its slot's stream, concurrently with the batches. Batched use submits
// all jobs at once; served use submits them as they arrive. Host inputs are
// copied at submit; meshes stay in the engine until taken.
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

namespace {

struct Job {
  gm_job desc{};
  std::vector<float> P, lattice;
  std::vector<uint32_t> T;
  int R = 0;
  int status = GM_OK;
  std::string err;
  bool done = false;
  double tSubmit = 0, tStart = 0, tDone = 0;
  std::vector<float> V, Nrm;
  std::vector<uint32_t> Tri;
  uint32_t nv = 0, nt = 0;
  int slot = -1;
};

// Pinned host buffer, grown on demand. Copies between pageable memory and
// the device go through the runtime's staging buffers, which the slots'
// threads would share; from pinned memory each slot's copies are plain DMA
// on its own stream.
template <class T>
struct Pinned {
  T* p = nullptr;
  size_t cap = 0;
  bool reserve(size_t n) {
    if (n <= cap) return true;
    if (p) hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = n + n / 4;
    if (hipHostMalloc((void**)&p, want * sizeof(T), hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return false;
    }
    cap = want;
    return true;
  }
  void release() {
    if (p) hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct Slot {
  gm_ctx* ctx = nullptr;
  int capR = 0;            // largest base resolution this slot has run
  uint64_t reserved = 0;   // device bytes the slot holds (estimate)
  std::thread th;
  Pinned<float> outV, outN;
  Pinned<uint32_t> outT;
  // hand-over to the batcher: the loaded job's structuring element, then the
  // outcome of its closing
  int se = 0;
  bool closed = false;
  int closeRc = 0;
  std::string closeErr;
};

// Device bytes of one context at base resolution R (the buffers alloc_grid,
// the mesh output and the radius field allocate), plus the mesh input.
// Jobs up to this base resolution are closed in batches; larger grids keep
// the GPU busy on their own.
constexpr int kBatchMaxRes = 192;

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
  b += (uint64_t)j.num_of_vertices * 24 + (uint64_t)j.num_of_faces * 12;
  return b + (16ull << 20);                       // tables, scans, slack
}

}  // namespace

struct gm_engine {
  int device = 0;
  uint64_t budget = 0;
  std::vector<Slot> slots;
  std::mutex m;
  std::condition_variable cvWork, cvDone;
  std::deque<Job*> queue;
  std::unordered_map<uint64_t, Job*> jobs;
  uint64_t nextId = 1;
  uint64_t reservedTotal = 0;
  int running = 0;
  bool stop = false;
  std::string err;
  // continuous batching: loaded slots wait in `ready`; the batcher thread
  // closes and meshes everything that is ready in one gm_batch_close_mesh
  // call (one launch per step for all of them) while the slots load the
  // next jobs
  gm_batch* batch = nullptr;
  std::thread batcher;
  std::mutex bm;
  std::condition_variable cvReady, cvClosed;
  std::vector<Slot*> ready;
  bool bstop = false;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // First queued job this slot can run within the budget (caller holds m).
  Job* pick(Slot& s) {
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      Job* j = *it;
      const int R = std::max(s.capR, j->R);
      const uint64_t need = std::max(s.reserved, context_bytes(R, j->desc.lattice_res, j->desc));
      const uint64_t grow = need - s.reserved;
      if (reservedTotal + grow > budget) continue;
      reservedTotal += grow;
      s.reserved = need;
      s.capR = R;
      queue.erase(it);
      return j;
    }
    return nullptr;
  }

  void run(Slot& s, Job* j) {
    gm_ctx* c = s.ctx;
    const gm_job& d = j->desc;
    auto step = [&](int rc) {
      if (rc != GM_OK && j->status == GM_OK) {
        j->status = rc;
        j->err = gm_last_error(c);
      }
      return rc == GM_OK;
    };
    auto oom = [&]() {
      j->status = GM_ERR_HIP;
      j->err = "GMorpho Error: pinned host buffer allocation failed";
    };
    // gm_load uploads through the context's own pinned staging
    if (!step(gm_load(c, j->P.data(), nullptr, d.num_of_vertices, j->T.data(), d.num_of_faces, d.base_res,
                      d.se_size)))
      return;
    if (!j->lattice.empty()) {
      if (!step(gm_scale_lattice_f32(c, j->lattice.data(), d.lattice_res))) return;
    } else if (!step(gm_scale_uniform(c, d.radius))) {
      return;
    }
    if (d.base_res <= kBatchMaxRes) {
      // small grid: closed with whatever else is ready
      std::unique_lock<std::mutex> lk(bm);
      s.se = d.se;
      s.closed = false;
      ready.push_back(&s);
      cvReady.notify_one();
      cvClosed.wait(lk, [&] { return s.closed; });
    } else {
      // large grid: enough work of its own, closed on the slot's stream
      // concurrently with the batches
      s.closeRc = gm_close(c, d.se);
      if (s.closeRc == GM_OK) s.closeRc = gm_extract_mesh(c);
      if (s.closeRc != GM_OK) s.closeErr = gm_last_error(c);
    }
    if (s.closeRc != GM_OK) {
      j->status = s.closeRc;
      j->err = s.closeErr;
      return;
    }
    if (!step(gm_mesh_counts(c, &j->nv, &j->nt))) return;
    const size_t nv3 = 3 * (size_t)j->nv, nt3 = 3 * (size_t)j->nt;
    if (!s.outV.reserve(nv3) || !s.outN.reserve(nv3) || !s.outT.reserve(nt3)) return oom();
    if (!step(gm_mesh_copy(c, s.outV.p, s.outN.p, s.outT.p))) return;
    j->V.assign(s.outV.p, s.outV.p + nv3);
    j->Nrm.assign(s.outN.p, s.outN.p + nv3);
    j->Tri.assign(s.outT.p, s.outT.p + nt3);
  }

  void batch_loop() {
    for (;;) {
      std::vector<Slot*> take;
      {
        std::unique_lock<std::mutex> lk(bm);
        cvReady.wait(lk, [&] { return bstop || !ready.empty(); });
        if (ready.empty()) return;  // stopping
        take.swap(ready);
      }
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
        int rc = gm_batch_close_mesh(batch, ctxs.data(), (int)ctxs.size(), se);
        std::vector<int> rcs(group.size(), rc);
        std::vector<std::string> msgs(group.size());
        if (rc != GM_OK) {
          // a job of the batch failed: each job's own status comes from
          // closing it alone (the loaded grids are unchanged by a closing)
          for (size_t i = 0; i < group.size(); ++i) {
            gm_ctx* c = group[i]->ctx;
            rcs[i] = gm_close(c, se);
            if (rcs[i] == GM_OK) rcs[i] = gm_extract_mesh(c);
            if (rcs[i] != GM_OK) msgs[i] = gm_last_error(c);
          }
        }
        std::lock_guard<std::mutex> lk(bm);
        for (size_t i = 0; i < group.size(); ++i) {
          group[i]->closeRc = rcs[i];
          group[i]->closeErr = msgs[i];
          group[i]->closed = true;
        }
      }
      cvClosed.notify_all();
    }
  }

  void worker(int idx) {
    Slot& s = slots[idx];
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      Job* j = nullptr;
      cvWork.wait(lk, [&] { return stop || (j = pick(s)) != nullptr; });
      if (!j) return;  // stopping
      j->slot = idx;
      j->tStart = now();
      ++running;
      lk.unlock();
      run(s, j);
      lk.lock();
      --running;
      j->tDone = now();
      j->done = true;
      cvDone.notify_all();
      cvWork.notify_all();
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
  if (!out || slots < 1 || slots > 64) return GM_ERR_ARG;
  *out = nullptr;
  gm_engine* e = new gm_engine();
  e->device = device;
  e->budget = hbm_budget_bytes;
  e->slots.resize(slots);
  for (auto& s : e->slots) {
    if (gm_create(device, &s.ctx) != GM_OK) {
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
  e->batcher = std::thread([e] { e->batch_loop(); });
  for (int i = 0; i < slots; ++i) e->slots[i].th = std::thread([e, i] { e->worker(i); });
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
  for (auto& s : e->slots)
    if (s.th.joinable()) s.th.join();
  // the workers are gone (the batcher served their last jobs): stop it
  {
    std::lock_guard<std::mutex> lk(e->bm);
    e->bstop = true;
  }
  e->cvReady.notify_all();
  if (e->batcher.joinable()) e->batcher.join();
  if (e->batch) gm_batch_destroy(e->batch);
  for (auto& s : e->slots) {
    if (s.ctx) gm_destroy(s.ctx);
    s.outV.release();
    s.outN.release();
    s.outT.release();
  }
  for (auto& kv : e->jobs) delete kv.second;
  delete e;
  return GM_OK;
}

const char* gm_engine_last_error(const gm_engine* e) { return e ? e->err.c_str() : ""; }

int gm_engine_submit(gm_engine* e, const gm_job* job, uint64_t* id) {
  if (!e || !job || !id) return GM_ERR_ARG;
  const gm_job& d = *job;
  if (!d.P || !d.T || d.num_of_vertices <= 0 || d.num_of_faces <= 0) return efail(e, GM_ERR_ARG, "empty mesh");
  if (d.base_res < 16 || d.base_res > 1024 || (d.base_res % 16) != 0)
    return efail(e, GM_ERR_UNSUPPORTED, "base resolution must be a multiple of 16 in [16,1024]");
  if (d.lattice && d.lattice_res < 2) return efail(e, GM_ERR_ARG, "lattice resolution below 2");
  Job* j = new Job();
  j->desc = d;
  j->R = d.base_res;
  j->P.assign(d.P, d.P + 3 * (size_t)d.num_of_vertices);
  j->T.assign(d.T, d.T + 3 * (size_t)d.num_of_faces);
  if (d.lattice) j->lattice.assign(d.lattice, d.lattice + (size_t)d.lattice_res * d.lattice_res * d.lattice_res);
  else j->desc.lattice_res = 0;
  j->desc.P = nullptr;
  j->desc.T = nullptr;
  j->desc.lattice = nullptr;
  if (context_bytes(j->R, j->desc.lattice_res, j->desc) > e->budget) {
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
    if (V) std::memcpy(V, j->V.data(), j->V.size() * 4);
    if (Nrm) std::memcpy(Nrm, j->Nrm.data(), j->Nrm.size() * 4);
    if (T) std::memcpy(T, j->Tri.data(), j->Tri.size() * 4);
  } else {
    e->err = j->err;
  }
  delete j;
  return rc;
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
