// Serving loop: DBA jobs (z-norm -> all-vs-all -> medoid -> DBA rounds to
// convergence, the performDBA pipeline for one cluster) submitted at any
// time, admitted in arrival order as soon as a preparation slot and enough
// device memory are free, and run concurrently.
//
// Continuous batching (the default): each of `slots` worker threads owns a
// context (a HIP stream) and prepares admitted jobs - upload, z-norm,
// all-vs-all, medoid (odba_job_prepare) - and hands each prepared job to
// the rounds engine of its value type and mode. An engine is a thread with
// its own context that runs one DBA round of every job it holds per step,
// packed into shared launches (odba_rounds_step); a job joins at the next
// step after its preparation and leaves when its own stop rule ends it, so
// jobs that arrive while others are on the device are batched with them.
// Preparations and the engines' rounds overlap on the device. A job's
// result does not depend on what runs beside it: every kernel is
// deterministic and each job's sums are formed in its own sequence order.
//
// With ODBA_SERVE_PACKED=0 in the environment at creation, each slot
// instead runs whole jobs (odba_perform_dba) one after another, and the
// slots' streams overlap.
//
// Admission compares an estimate of the job's device working set
// (resident sequences, distances, edge scratch, run records, centroids)
// with what the jobs already admitted leave of the HBM budget. A prepared
// job then counts what it keeps (sequences and run records) and its step
// words (a buffer of its own for the job's life, sized once the medoid is
// known), which its preparing slot allocates as soon as they fit the
// budget (at once when no prepared job holds words), so the engines never
// allocate or free them on their steps.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/odba/odba.h"

namespace {

size_t elem_size(int dtype) {
  switch (dtype) {
    case ODBA_I32: case ODBA_U32: case ODBA_F32: return 4;
    case ODBA_U64: case ODBA_F64: return 8;
  }
  return 0;
}

long long padded(long long L) { return ((L + 63) / 64) * 64 + 1024; }

struct Job {
  uint64_t id = 0;
  int dtype = 0;
  std::vector<std::vector<unsigned char>> data;
  std::vector<uint64_t> lens;
  int os = 0, oe = 0, norm = 1, max_rounds = 250;
  uint64_t footprint = 0;
  bool worded = false;  // step words counted in the server's in_use
  int state = ODBA_JOB_QUEUED;
  int rc = 0;
  std::string err;
  double t_submit = 0, t_admit = 0, t_done = 0;
  std::vector<unsigned char> centroid;
  uint64_t clen = 0;
  int medoid = -1, rounds = 0;
  odba_job* handle = nullptr;  // prepared job while its rounds run
};

// Rounds engine of one (value type, open start, open end) class.
struct Engine {
  odba_ctx* ctx = nullptr;
  odba_rounds* eng = nullptr;
  std::thread th;
  std::deque<Job*> pending;   // prepared, not yet added
  int held = 0;               // jobs added and not yet returned
};

}  // namespace

struct odba_server {
  int device = 0;
  uint64_t budget = 0;
  int slots = 1;
  int cus = 256;
  bool packed = true;
  std::vector<odba_ctx*> ctxs;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_admit, cv_done, cv_engine, cv_reap;
  std::deque<Job*> queue;
  std::map<uint64_t, std::unique_ptr<Job>> jobs;
  std::map<int, std::unique_ptr<Engine>> engines;
  std::deque<odba_job*> reap;
  std::thread reaper;
  uint64_t in_use = 0;
  int worded = 0;         // jobs whose step words are counted in in_use
  int running = 0;
  bool stop = false;
  uint64_t next_id = 1;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::string err;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // Device bytes one job holds while it is prepared or runs, apart from
  // step words (a prepared job's own buffer; with whole-job slots, the
  // slot's share). With `held`, what a prepared job keeps: its sequences
  // and run records.
  uint64_t footprint(const Job& j, bool held = false) const {
    const size_t es = elem_size(j.dtype);
    const long long n = (long long)j.lens.size();
    long long seq = 0, maxlen = 0, sumlen = 0;
    for (uint64_t L : j.lens) {
      seq += padded((long long)L);
      maxlen = std::max<long long>(maxlen, (long long)L);
      sumlen += (long long)L;
    }
    const long long npairs = n * (n - 1) / 2;
    const long long waves = std::min<long long>((npairs + 3) / 4, (long long)cus * 8) * 4;
    const uint64_t edge = (uint64_t)waves * 2 * (maxlen + 256) * es;
    const uint64_t dists = (uint64_t)npairs * es;
    const uint64_t runs = (uint64_t)n * maxlen * 16 + (uint64_t)padded(maxlen) * es * 4;
    uint64_t words = 0;
    if (!packed) {
      words = (uint64_t)sumlen * (uint64_t)maxlen / 4 + (uint64_t)n * 64 * 4 * (maxlen / 1024 + 1);
      words = std::min(words, budget / (uint64_t)slots / 4);
    }
    if (held) return (uint64_t)seq * es * 2 + runs;
    return (uint64_t)seq * es * 2 + dists + edge + runs + words;
  }

  // Whole-job slots bound the jobs in flight; with packed rounds they bound
  // the preparations only, and every prepared job joins its engine.
  bool can_admit() const {
    if (queue.empty() || (packed ? prep_slots_busy >= slots : running >= slots)) return false;
    const Job* j = queue.front();
    return running == 0 || in_use + j->footprint <= budget;
  }

  static int class_key(const Job& j) { return j.dtype * 4 + (j.os ? 2 : 0) + (j.oe ? 1 : 0); }

  // Records a finished job's outcome (lock held).
  void complete(Job* j, int rc, const std::string& msg) {
    in_use -= j->footprint;
    if (j->worded) worded--;
    j->worded = false;
    j->rc = rc;
    if (rc) j->err = msg;
    j->data.clear();
    j->data.shrink_to_fit();
    j->t_done = now();
    j->state = rc ? ODBA_JOB_FAILED : ODBA_JOB_DONE;
  }

  // Whole-job slots: odba_perform_dba per admitted job.
  void worker_whole(int slot) {
    odba_ctx* c = ctxs[slot];
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_admit.wait(lk, [&] { return stop || can_admit(); });
        if (stop) return;
        j = queue.front();
        queue.pop_front();
        in_use += j->footprint;
        running++;
        j->state = ODBA_JOB_RUNNING;
        j->t_admit = now();
      }
      const int n = (int)j->lens.size();
      std::vector<const void*> ptrs(n);
      for (int i = 0; i < n; i++) ptrs[i] = j->data[i].data();
      uint64_t maxlen = 0;
      for (uint64_t L : j->lens) maxlen = std::max(maxlen, L);
      std::vector<unsigned char> cen(maxlen * elem_size(j->dtype));
      uint64_t clen = 0;
      int med = -1, rounds = 0;
      const int rc = odba_perform_dba(c, j->dtype, ptrs.data(), j->lens.data(), n, j->os, j->oe, j->norm,
                                      j->max_rounds, ODBA_PATH_FULL, 0, cen.data(), &clen, &med, nullptr, &rounds,
                                      nullptr);
      {
        std::lock_guard<std::mutex> lk(mu);
        running--;
        j->centroid.assign(cen.begin(), cen.begin() + clen * elem_size(j->dtype));
        j->clen = clen;
        j->medoid = med;
        j->rounds = rounds;
        complete(j, rc, rc ? odba_last_error(c) : "");
      }
      cv_done.notify_all();
      cv_admit.notify_all();
    }
  }

  // Packed rounds: a slot prepares admitted jobs and hands them on.
  void worker_prepare(int slot) {
    odba_ctx* c = ctxs[slot];
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_admit.wait(lk, [&] { return stop || can_admit(); });
        if (stop) return;
        j = queue.front();
        queue.pop_front();
        in_use += j->footprint;
        running++;
        prep_slots_busy++;
        j->state = ODBA_JOB_RUNNING;
        j->t_admit = now();
      }
      const int n = (int)j->lens.size();
      std::vector<const void*> ptrs(n);
      for (int i = 0; i < n; i++) ptrs[i] = j->data[i].data();
      odba_job* h = nullptr;
      int rc = odba_job_prepare(c, j->dtype, ptrs.data(), j->lens.data(), n, j->os, j->oe, j->norm, j->max_rounds,
                                &h);
      std::string msg = rc ? odba_last_error(c) : "";
      uint64_t wb = 0;
      if (!rc && (rc = odba_job_words(h, 0, &wb))) msg = "cannot size the job's step words";
      if (!rc) {
        std::unique_lock<std::mutex> lk(mu);
        in_use -= j->footprint;
        j->footprint = footprint(*j, true);
        in_use += j->footprint;
        cv_admit.wait(lk, [&] { return stop || worded == 0 || in_use + wb <= budget; });
        in_use += wb;
        j->footprint += wb;
        j->worded = true;
        worded++;
      }
      if (!rc && (rc = odba_job_words(h, 1, &wb))) msg = "cannot allocate the job's step words";
      {
        std::lock_guard<std::mutex> lk(mu);
        prep_slots_busy--;
        if (rc) {
          running--;
          if (h) reap.push_back(h);
          complete(j, rc, msg);
        } else {
          j->handle = h;
          j->data.clear();
          j->data.shrink_to_fit();
          Engine* e = engine_for(*j);
          if (!e) {
            running--;
            reap.push_back(h);
            j->handle = nullptr;
            complete(j, ODBA_E_NOMEM, err);
          } else {
            e->pending.push_back(j);
          }
        }
      }
      cv_engine.notify_all();
      cv_reap.notify_all();
      cv_done.notify_all();
      cv_admit.notify_all();
    }
  }
  int prep_slots_busy = 0;

  // The engine of a job's class, created on first use (lock held).
  Engine* engine_for(const Job& j) {
    const int key = class_key(j);
    auto it = engines.find(key);
    if (it != engines.end()) return it->second.get();
    auto e = std::make_unique<Engine>();
    if (odba_ctx_create(device, &e->ctx)) {
      err = "cannot create a rounds-engine context";
      return nullptr;
    }
    odba_ctx_set_graphs(e->ctx, 0);
    if (odba_rounds_create(e->ctx, j.dtype, j.os, j.oe, &e->eng)) {
      err = odba_last_error(e->ctx);
      odba_ctx_destroy(e->ctx);
      return nullptr;
    }
    Engine* ep = e.get();
    e->th = std::thread([this, ep] { engine_loop(ep); });
    engines[key] = std::move(e);
    return ep;
  }

  void engine_loop(Engine* e) {
    std::vector<odba_job*> fin(64);
    std::map<odba_job*, Job*> owner;
    for (;;) {
      std::vector<Job*> add;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_engine.wait(lk, [&] { return (stop && preps_done) || !e->pending.empty() || e->held > 0; });
        if (stop && preps_done && e->held == 0 && e->pending.empty()) return;
        add.assign(e->pending.begin(), e->pending.end());
        e->pending.clear();
      }
      for (Job* j : add) {
        const int rc = odba_rounds_add(e->eng, j->handle);
        std::lock_guard<std::mutex> lk(mu);
        if (rc) {
          running--;
          reap.push_back(j->handle);
          j->handle = nullptr;
          complete(j, rc, odba_last_error(e->ctx));
        } else {
          owner[j->handle] = j;
          e->held++;
        }
      }
      int nf = 0, na = 0;
      odba_rounds_step(e->eng, fin.data(), (int)fin.size(), &nf, &na);
      if (nf == 0) continue;
      {
        std::lock_guard<std::mutex> lk(mu);
        for (int k = 0; k < nf; k++) {
          odba_job* h = fin[k];
          Job* j = owner[h];
          owner.erase(h);
          e->held--;
          running--;
          int rc = ODBA_OK;
          uint64_t clen = 0;
          odba_job_result(h, nullptr, 0, &clen, &j->medoid, &j->rounds, &rc);
          j->centroid.resize(clen * elem_size(j->dtype));
          odba_job_result(h, j->centroid.data(), clen, &j->clen, nullptr, nullptr, nullptr);
          j->handle = nullptr;
          reap.push_back(h);
          complete(j, rc, rc ? odba_last_error(e->ctx) : "");
        }
      }
      cv_reap.notify_all();
      cv_done.notify_all();
      cv_admit.notify_all();
    }
  }

  // Device memory of finished jobs is freed here: hipFree waits for the
  // work queued on the device, which would stall an engine or a slot.
  void reap_loop() {
    for (;;) {
      odba_job* h = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_reap.wait(lk, [&] { return !reap.empty() || (stop && workers_done); });
        if (reap.empty()) return;
        h = reap.front();
        reap.pop_front();
      }
      odba_job_destroy(h);
    }
  }
  bool workers_done = false;
  bool preps_done = false;
};

extern "C" {

int odba_server_create(int device, uint64_t hbm_budget, int slots, odba_server** out) {
  if (!out || slots < 1 || slots > 32) return ODBA_E_INVALID;
  auto* s = new odba_server();
  s->device = device;
  s->slots = slots;
  const char* pk = std::getenv("ODBA_SERVE_PACKED");
  s->packed = !(pk && !std::strcmp(pk, "0"));
  for (int k = 0; k < slots; k++) {
    odba_ctx* c = nullptr;
    const int rc = odba_ctx_create(device, &c);
    if (rc) {
      for (odba_ctx* x : s->ctxs) odba_ctx_destroy(x);
      delete s;
      return rc;
    }
    s->ctxs.push_back(c);
  }
  int cus = 0, clk = 0;
  uint64_t mem = 0;
  char name[64];
  odba_device_info(s->ctxs[0], &cus, &clk, &mem, name, sizeof(name));
  s->cus = cus > 0 ? cus : 256;
  s->budget = hbm_budget ? std::min<uint64_t>(hbm_budget, mem) : mem / 10 * 8;
  for (odba_ctx* c : s->ctxs) {
    odba_ctx_set_mem_cap(c, s->budget / (uint64_t)slots);
    odba_ctx_set_graphs(c, 0);  // graph replays serialise the slots' streams (measured)
  }
  if (s->packed) {
    s->reaper = std::thread([s] { s->reap_loop(); });
    for (int k = 0; k < slots; k++) s->workers.emplace_back([s, k] { s->worker_prepare(k); });
  } else {
    for (int k = 0; k < slots; k++) s->workers.emplace_back([s, k] { s->worker_whole(k); });
  }
  *out = s;
  return ODBA_OK;
}

int odba_server_destroy(odba_server* s) {
  if (!s) return ODBA_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (Job* j : s->queue) {
      j->state = ODBA_JOB_FAILED;
      j->rc = ODBA_E_BUSY;
    }
    s->queue.clear();
  }
  s->cv_admit.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->workers) t.join();  // preparations in flight are handed to their engines
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->preps_done = true;
  }
  s->cv_engine.notify_all();
  for (auto& kv : s->engines) {  // engines finish the jobs they hold
    Engine* e = kv.second.get();
    if (e->th.joinable()) e->th.join();
    odba_rounds_destroy(e->eng);
    odba_ctx_destroy(e->ctx);
  }
  if (s->reaper.joinable()) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->workers_done = true;
    }
    s->cv_reap.notify_all();
    s->reaper.join();
  }
  for (odba_ctx* c : s->ctxs) odba_ctx_destroy(c);
  delete s;
  return ODBA_OK;
}

int odba_server_submit(odba_server* s, int dtype, const void* const* seqs, const uint64_t* lens, int n,
                       int use_open_start, int use_open_end, int norm, int max_rounds, uint64_t* job_id) {
  if (!s || !seqs || !lens || n < 2 || !job_id || max_rounds < 1) return ODBA_E_INVALID;
  const size_t es = elem_size(dtype);
  if (!es) return ODBA_E_INVALID;
  auto j = std::make_unique<Job>();
  j->dtype = dtype;
  j->lens.assign(lens, lens + n);
  j->data.resize(n);
  for (int i = 0; i < n; i++) {
    if (!seqs[i] || lens[i] == 0) return ODBA_E_INVALID;
    const unsigned char* p = (const unsigned char*)seqs[i];
    j->data[i].assign(p, p + lens[i] * es);
  }
  j->os = use_open_start;
  j->oe = use_open_end;
  j->norm = norm;
  j->max_rounds = max_rounds;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return ODBA_E_BUSY;
    j->id = s->next_id++;
    j->footprint = s->footprint(*j);
    j->t_submit = s->now();
    *job_id = j->id;
    s->queue.push_back(j.get());
    s->jobs[j->id] = std::move(j);
  }
  s->cv_admit.notify_all();
  return ODBA_OK;
}

int odba_server_wait(odba_server* s, uint64_t job_id, double timeout_s, int* state) {
  if (!s) return ODBA_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return ODBA_E_INVALID;
  Job* j = it->second.get();
  auto done = [&] { return j->state == ODBA_JOB_DONE || j->state == ODBA_JOB_FAILED; };
  if (timeout_s < 0) s->cv_done.wait(lk, done);
  else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  if (state) *state = j->state;
  return ODBA_OK;
}

int odba_server_result(odba_server* s, uint64_t job_id, void* centroid_out, uint64_t capacity,
                       uint64_t* centroid_len, int* medoid, int* rounds, double* times, int* job_rc) {
  if (!s) return ODBA_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return ODBA_E_INVALID;
  Job* j = it->second.get();
  if (j->state != ODBA_JOB_DONE && j->state != ODBA_JOB_FAILED) return ODBA_E_BUSY;
  if (centroid_out) {
    if (capacity < j->clen) return ODBA_E_INVALID;
    std::memcpy(centroid_out, j->centroid.data(), j->centroid.size());
  }
  if (centroid_len) *centroid_len = j->clen;
  if (medoid) *medoid = j->medoid;
  if (rounds) *rounds = j->rounds;
  if (times) {
    times[0] = j->t_submit;
    times[1] = j->t_admit;
    times[2] = j->t_done;
  }
  if (job_rc) *job_rc = j->rc;
  return ODBA_OK;
}

int odba_server_release(odba_server* s, uint64_t job_id) {
  if (!s) return ODBA_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return ODBA_E_INVALID;
  if (it->second->state != ODBA_JOB_DONE && it->second->state != ODBA_JOB_FAILED) return ODBA_E_BUSY;
  s->jobs.erase(it);
  return ODBA_OK;
}

int odba_server_stats(odba_server* s, uint64_t* bytes_in_use, uint64_t* budget, int* running, int* queued) {
  if (!s) return ODBA_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (bytes_in_use) *bytes_in_use = s->in_use;
  if (budget) *budget = s->budget;
  if (running) *running = s->running;
  if (queued) *queued = (int)s->queue.size();
  return ODBA_OK;
}

}  // extern "C"
