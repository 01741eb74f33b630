// mc_server.cpp - continuous-batching step server, built only on the public
// C ABI (minecu_hip.h): a loop thread merges the pending step requests of all
// admitted env groups into one mc_batch_step launch per free stream, retires
// launches as their completion events fire, and records per-request latency.
//
// Groups stay resident on the device for their whole session.  Admission is
// by an HBM budget: a group whose bytes do not fit waits (FIFO) until earlier
// groups are removed.  A group has at most one request on the device at a
// time (its steps are sequential); requests that arrive meanwhile queue and
// join the next launch, which can run on another stream while the current one
// drains.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <iterator>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "minecu_hip.h"

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

thread_local char g_serr[256] = "";

int sfail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_serr, sizeof g_serr, fmt, ap);
  va_end(ap);
  return code;
}

struct Req {
  int64_t id;
  int gid;
  const void* actions;
  int kind;
  const float* look;
  uint8_t* dones;
  int64_t t_submit;
};

enum GroupState { kWaiting = 0, kAdmitted = 1, kRemoved = 2 };

struct Group {
  mc_env* env;
  int64_t bytes;
  int batch;
  double cost;     // relative work of one step (batch x per-env rays and voxels)
  int state;
  bool busy;       // a request of this group is on the device
  int pending;     // queued requests
  // Closed-loop replay: after each completion, submit the next of `rep_count`
  // steps whose actions start at rep_base (+ k * rep_stride bytes).
  const char* rep_base;
  int64_t rep_stride;
  int64_t rep_count;
  int64_t rep_next;
  int rep_kind;
  const char* rep_look;
  int64_t rep_look_stride;
  uint8_t* rep_dones;
};

struct Launch {
  hipEvent_t ev;
  int stream;
  std::vector<Req> reqs;
  int64_t t_launch;
};

}  // namespace

struct mc_server {
  std::mutex mu;
  std::condition_variable cv;
  int64_t budget = 0;
  int64_t used = 0;
  double admitted_cost = 0.0;  // summed step cost of the admitted groups
  std::vector<Group> groups;
  std::deque<int> waiting;
  std::deque<Req> pending;
  // Planned arrivals (mc_server_schedule), in time order: released into
  // `pending` by the loop thread once their time has come.
  std::vector<Req> sched;
  size_t sched_next = 0;
  std::deque<Launch> inflight;
  // Requests taken from `pending` whose launch is being issued (outside the
  // lock) and is not in `inflight` yet; drain waits for them too.
  int launching = 0;
  std::vector<hipStream_t> streams;
  std::vector<char> stream_busy;
  std::vector<hipEvent_t> events;
  mc_batch* batch = nullptr;
  std::thread th;
  bool running = false;
  int device = 0;
  std::vector<mc_server_record> done;
  int64_t next_id = 1;
  int64_t launches = 0, env_steps = 0, grouped = 0, failed = 0;
  std::string last_err;

  void admit_waiting() {
    while (!waiting.empty()) {
      Group& g = groups[waiting.front()];
      if (used + g.bytes > budget) break;
      used += g.bytes;
      admitted_cost += g.cost;
      g.state = kAdmitted;
      waiting.pop_front();
    }
  }

  void retire_locked(int64_t t) {
    for (auto it = inflight.begin(); it != inflight.end();) {
      if (hipEventQuery(it->ev) != hipSuccess) {
        ++it;
        continue;
      }
      for (const Req& r : it->reqs) {
        done.push_back(mc_server_record{r.id, r.gid, r.t_submit, it->t_launch, t});
        Group& g = groups[r.gid];
        g.busy = false;
        if (g.rep_base && g.rep_next < g.rep_count) {
          const int64_t k = g.rep_next++;
          pending.push_back(Req{next_id++, r.gid, g.rep_base + k * g.rep_stride, g.rep_kind,
                                g.rep_look ? (const float*)(g.rep_look + k * g.rep_look_stride)
                                           : nullptr,
                                g.rep_dones, t});
          g.pending++;
        }
      }
      stream_busy[it->stream] = 0;
      events.push_back(it->ev);
      it = inflight.erase(it);
    }
  }

  // Queue the planned arrivals whose time has come (t_submit = arrival).
  void release_scheduled_locked(int64_t t) {
    while (sched_next < sched.size() && sched[sched_next].t_submit <= t) {
      const Req& r = sched[sched_next++];
      if (groups[r.gid].state != kAdmitted) {
        failed++;
        last_err = "scheduled request for a group that is no longer admitted";
        done.push_back(mc_server_record{r.id, r.gid, r.t_submit, -1, -1});
        continue;
      }
      pending.push_back(r);
      groups[r.gid].pending++;
    }
    if (sched_next == sched.size() && !sched.empty()) {
      sched.clear();
      sched_next = 0;
    }
  }

  void loop() {
    (void)hipSetDevice(device);
    std::vector<Req> take;
    std::vector<mc_env*> envs;
    std::vector<const void*> acts;
    std::vector<int> kinds;
    std::vector<const float*> looks;
    std::vector<uint8_t*> dns;
    for (;;) {
      int stream = -1;
      hipEvent_t ev = nullptr;
      take.clear();
      {
        std::unique_lock<std::mutex> lk(mu);
        const int64_t t_now = now_ns();
        retire_locked(t_now);
        if (running) release_scheduled_locked(t_now);
        if (!running && pending.empty() && inflight.empty()) break;
        for (size_t i = 0; i < streams.size(); ++i)
          if (!stream_busy[i]) {
            stream = (int)i;
            break;
          }
        if (stream >= 0 && !events.empty()) {
          // With several streams a launch takes at most its share of the
          // admitted work, so when every group is waiting (a saturated, closed
          // loop) the groups form cohorts whose launches queue behind each
          // other: the device starts the next one while the host retires the
          // last and gathers its successors.
          const double cap = streams.size() > 1 ? admitted_cost / (double)streams.size() : 1e300;
          double taken = 0.0;
          for (auto it = pending.begin();
               it != pending.end() && (int)take.size() < 64;) {
            Group& g = groups[it->gid];
            bool dup = false;
            for (const Req& q : take) dup |= (q.gid == it->gid);
            if (g.state == kAdmitted && !g.busy && !dup) {
              if (!take.empty() && taken + g.cost > cap * 1.0001) break;
              taken += g.cost;
              take.push_back(*it);
              g.busy = true;
              g.pending--;
              it = pending.erase(it);
            } else {
              ++it;
            }
          }
        }
        launching += (int)take.size();
        if (take.empty()) {
          // Sleep until a submit, or until shortly before the next planned
          // arrival (then poll, as while launches are in flight).
          int64_t wait_ns = 2000000;
          if (sched_next < sched.size())
            wait_ns = std::min<int64_t>(wait_ns, sched[sched_next].t_submit - now_ns() - 150000);
          if (!inflight.empty() || wait_ns <= 0) {
            lk.unlock();
            std::this_thread::yield();
          } else {
            cv.wait_for(lk, std::chrono::nanoseconds(wait_ns));
          }
          continue;
        }
        stream_busy[stream] = 1;
        ev = events.back();
        events.pop_back();
      }
      envs.clear();
      acts.clear();
      kinds.clear();
      looks.clear();
      dns.clear();
      int64_t steps = 0;
      for (const Req& r : take) {
        envs.push_back(groups[r.gid].env);
        acts.push_back(r.actions);
        kinds.push_back(r.kind);
        looks.push_back(r.look);
        dns.push_back(r.dones);
        steps += groups[r.gid].batch;
      }
      const int64_t t_launch = now_ns();
      int st = mc_batch_step(batch, envs.data(), acts.data(), kinds.data(), looks.data(),
                             dns.data(), (int)take.size(), streams[stream]);
      if (st == MC_OK && hipEventRecord(ev, streams[stream]) != hipSuccess) st = MC_E_HIP;
      std::lock_guard<std::mutex> lk(mu);
      launching -= (int)take.size();
      if (st != MC_OK) {
        last_err = mc_last_error();
        failed += (int64_t)take.size();
        for (const Req& r : take) {
          groups[r.gid].busy = false;
          done.push_back(mc_server_record{r.id, r.gid, r.t_submit, t_launch, -1});
        }
        stream_busy[stream] = 0;
        events.push_back(ev);
        continue;
      }
      launches++;
      grouped += (int64_t)take.size();
      env_steps += steps;
      inflight.push_back(Launch{ev, stream, take, t_launch});
    }
  }
};

extern "C" {

const char* mc_server_last_error(mc_server* s) {
  if (!s) return g_serr;
  std::lock_guard<std::mutex> lk(s->mu);
  return s->last_err.empty() ? g_serr : s->last_err.c_str();
}

int mc_server_create(int64_t hbm_budget_bytes, int n_streams, mc_server** out) {
  if (!out) return sfail(MC_E_INVALID, "out is NULL");
  *out = nullptr;
  if (hbm_budget_bytes <= 0) return sfail(MC_E_INVALID, "budget must be positive");
  if (n_streams < 1 || n_streams > 8) return sfail(MC_E_INVALID, "n_streams must be in [1, 8]");
  mc_server* s = new (std::nothrow) mc_server;
  if (!s) return sfail(MC_E_INVALID, "out of host memory");
  s->budget = hbm_budget_bytes;
  hipError_t e = hipGetDevice(&s->device);
  for (int i = 0; i < n_streams && e == hipSuccess; ++i) {
    hipStream_t st = nullptr;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) {
      s->streams.push_back(st);
      s->stream_busy.push_back(0);
    }
  }
  for (int i = 0; i < 2 * n_streams && e == hipSuccess; ++i) {
    hipEvent_t ev = nullptr;
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) s->events.push_back(ev);
  }
  int r = (e == hipSuccess) ? mc_batch_create(&s->batch) : MC_E_HIP;
  if (e != hipSuccess || r != MC_OK) {
    mc_server_destroy(s);
    return sfail(MC_E_HIP, "server create failed");
  }
  s->running = true;
  s->th = std::thread([s] { s->loop(); });
  *out = s;
  return MC_OK;
}

int mc_server_destroy(mc_server* s) {
  if (!s) return MC_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->running = false;
    // Requests that were never launched are dropped.
    s->pending.clear();
  }
  s->cv.notify_all();
  if (s->th.joinable()) s->th.join();
  for (hipEvent_t ev : s->events) (void)hipEventDestroy(ev);
  for (auto& l : s->inflight) (void)hipEventDestroy(l.ev);
  for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
  if (s->batch) mc_batch_destroy(s->batch);
  delete s;
  return MC_OK;
}

int mc_server_add_group(mc_server* s, mc_env* env, int64_t bytes, int batch_size, int* gid) {
  if (!s || !env || !gid || bytes < 0 || batch_size < 1) return sfail(MC_E_INVALID, "bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (bytes > s->budget) return sfail(MC_E_INVALID, "group needs more than the whole budget");
  mc_config c{};
  if (mc_env_get_config(env, &c) != MC_OK) return sfail(MC_E_INVALID, "bad env handle");
  Group g{};
  g.env = env;
  g.bytes = bytes;
  g.batch = batch_size;
  // Step time per env on MI355X fits rays x world height (DDA length) plus
  // twice the world's voxels (staging, scan, regeneration) to within ~15%
  // over 16^3/32^3 worlds and 32^2/64^2 images.
  g.cost = (double)batch_size * ((double)c.width * c.height * c.world_y +
                                 2.0 * c.world_x * c.world_y * c.world_z);
  g.state = kWaiting;
  *gid = (int)s->groups.size();
  s->groups.push_back(g);
  s->waiting.push_back(*gid);
  s->admit_waiting();
  return MC_OK;
}

int mc_server_group_state(mc_server* s, int gid, int* state) {
  if (!s || !state) return sfail(MC_E_INVALID, "bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (gid < 0 || gid >= (int)s->groups.size()) return sfail(MC_E_INVALID, "no group %d", gid);
  *state = s->groups[gid].state;
  return MC_OK;
}

int mc_server_remove_group(mc_server* s, int gid) {
  if (!s) return sfail(MC_E_INVALID, "server is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  if (gid < 0 || gid >= (int)s->groups.size()) return sfail(MC_E_INVALID, "no group %d", gid);
  Group& g = s->groups[gid];
  if (g.busy || g.pending > 0) return sfail(MC_E_INVALID, "group %d has requests in flight", gid);
  if (g.state == kAdmitted) {
    s->used -= g.bytes;
    s->admitted_cost -= g.cost;
  }
  if (g.state == kWaiting)
    s->waiting.erase(std::remove(s->waiting.begin(), s->waiting.end(), gid), s->waiting.end());
  g.state = kRemoved;
  g.rep_base = nullptr;
  s->admit_waiting();
  return MC_OK;
}

int mc_server_submit(mc_server* s, int gid, const void* actions, int action_kind,
                     const float* look, uint8_t* dones, int64_t* req_id) {
  if (!s || !actions) return sfail(MC_E_INVALID, "bad argument");
  if (action_kind == MC_ACT_BUTTONS && !look) return sfail(MC_E_INVALID, "look is NULL");
  int64_t id;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (gid < 0 || gid >= (int)s->groups.size()) return sfail(MC_E_INVALID, "no group %d", gid);
    if (s->groups[gid].state != kAdmitted)
      return sfail(MC_E_INVALID, "group %d is not admitted", gid);
    id = s->next_id++;
    s->pending.push_back(Req{id, gid, actions, action_kind, look, dones, now_ns()});
    s->groups[gid].pending++;
  }
  s->cv.notify_one();
  if (req_id) *req_id = id;
  return MC_OK;
}

int mc_server_set_replay(mc_server* s, int gid, const void* actions_base, int64_t stride_bytes,
                         int64_t count, int action_kind, const float* look_base,
                         int64_t look_stride_bytes, uint8_t* dones) {
  if (!s || !actions_base || count < 0) return sfail(MC_E_INVALID, "bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (gid < 0 || gid >= (int)s->groups.size()) return sfail(MC_E_INVALID, "no group %d", gid);
  Group& g = s->groups[gid];
  g.rep_base = (const char*)actions_base;
  g.rep_stride = stride_bytes;
  g.rep_count = count;
  g.rep_next = 0;
  g.rep_kind = action_kind;
  g.rep_look = (const char*)look_base;
  g.rep_look_stride = look_stride_bytes;
  g.rep_dones = dones;
  return MC_OK;
}

int mc_server_start_replay(mc_server* s, int gid) {
  if (!s) return sfail(MC_E_INVALID, "server is NULL");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (gid < 0 || gid >= (int)s->groups.size()) return sfail(MC_E_INVALID, "no group %d", gid);
    Group& g = s->groups[gid];
    if (!g.rep_base || g.rep_next >= g.rep_count || g.state != kAdmitted)
      return sfail(MC_E_INVALID, "group %d has no replay to start", gid);
    const int64_t k = g.rep_next++;
    s->pending.push_back(Req{s->next_id++, gid, g.rep_base + k * g.rep_stride, g.rep_kind,
                             g.rep_look ? (const float*)(g.rep_look + k * g.rep_look_stride) : nullptr,
                             g.rep_dones, now_ns()});
    g.pending++;
  }
  s->cv.notify_one();
  return MC_OK;
}

int mc_server_schedule(mc_server* s, const int* gids, const void* const* actions,
                       int action_kind, const int64_t* t_offset_ns, int64_t n) {
  if (!s || n < 0 || (n > 0 && (!gids || !actions || !t_offset_ns)))
    return sfail(MC_E_INVALID, "bad argument");
  if (action_kind != MC_ACT_DISCRETE && action_kind != MC_ACT_BUTTONS)
    return sfail(MC_E_INVALID, "bad action kind %d", action_kind);
  if (action_kind == MC_ACT_BUTTONS) return sfail(MC_E_INVALID, "scheduled requests carry no look");
  for (int64_t k = 0; k < n; ++k) {
    if (!actions[k]) return sfail(MC_E_INVALID, "request %lld: actions is NULL", (long long)k);
    if (t_offset_ns[k] < 0 || (k > 0 && t_offset_ns[k] < t_offset_ns[k - 1]))
      return sfail(MC_E_INVALID, "arrival offsets must be non-negative and non-decreasing");
  }
  // Built outside the lock with relative times; the clock starts once the
  // list is ready, so building a long schedule delays no arrival.
  std::vector<Req> add;
  add.reserve((size_t)n);
  for (int64_t k = 0; k < n; ++k)
    add.push_back(Req{0, gids[k], actions[k], action_kind, nullptr, nullptr, t_offset_ns[k]});
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (int64_t k = 0; k < n; ++k)
      if (gids[k] < 0 || gids[k] >= (int)s->groups.size() || s->groups[gids[k]].state != kAdmitted)
        return sfail(MC_E_INVALID, "request %lld: group %d is not admitted", (long long)k, gids[k]);
    // Completion records are appended under the lock: reserve now, so no
    // reallocation of a long record list stalls the loop mid-schedule.
    s->done.reserve(s->done.size() + (s->sched.size() - s->sched_next) + add.size());
    const int64_t t0 = now_ns();
    for (Req& r : add) {
      r.id = s->next_id++;
      r.t_submit += t0;
    }
    if (s->sched_next >= s->sched.size()) {
      s->sched.swap(add);
    } else {
      // Merge with what is left of a running schedule, keeping time order.
      std::vector<Req> merged;
      merged.reserve(s->sched.size() - s->sched_next + add.size());
      std::merge(s->sched.begin() + (int64_t)s->sched_next, s->sched.end(), add.begin(), add.end(),
                 std::back_inserter(merged),
                 [](const Req& a, const Req& b) { return a.t_submit < b.t_submit; });
      s->sched.swap(merged);
    }
    s->sched_next = 0;
  }
  s->cv.notify_one();
  return MC_OK;
}

int mc_server_drain(mc_server* s) {
  if (!s) return sfail(MC_E_INVALID, "server is NULL");
  for (;;) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      bool replay_left = false;
      for (const Group& g : s->groups) replay_left |= (g.rep_base && g.rep_next < g.rep_count);
      if (s->pending.empty() && s->inflight.empty() && s->launching == 0 && !replay_left &&
          s->sched_next >= s->sched.size())
        break;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(100));
  }
  std::lock_guard<std::mutex> lk(s->mu);
  return s->failed ? sfail(MC_E_HIP, "%lld requests failed: %s", (long long)s->failed,
                           s->last_err.c_str())
                   : MC_OK;
}

int mc_server_take_records(mc_server* s, mc_server_record* out, int max, int* n) {
  if (!s || !n || (max > 0 && !out)) return sfail(MC_E_INVALID, "bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  const int k = std::min<int>(max, (int)s->done.size());
  std::copy(s->done.begin(), s->done.begin() + k, out);
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return MC_OK;
}

int mc_server_stats(mc_server* s, int64_t* launches, int64_t* env_steps, int64_t* requests,
                    int64_t* used_bytes) {
  if (!s) return sfail(MC_E_INVALID, "server is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  if (launches) *launches = s->launches;
  if (env_steps) *env_steps = s->env_steps;
  if (requests) *requests = s->grouped;
  if (used_bytes) *used_bytes = s->used;
  return MC_OK;
}

int mc_server_autotune(mc_server* s, int steps, void* stream) {
  if (!s) return sfail(MC_E_INVALID, "server is NULL");
  std::vector<mc_env*> envs;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (const Group& g : s->groups)
      if (g.state == kAdmitted && !g.busy && g.pending == 0) envs.push_back(g.env);
  }
  if (envs.empty()) return MC_OK;
  return mc_batch_autotune(s->batch, envs.data(), (int)envs.size(), steps, stream);
}

}  // extern "C"
