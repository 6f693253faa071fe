// api_serve.cpp -- request server over nckde sessions (continuous batching).
//
// Worker threads take requests in arrival order, with backfill: a request
// whose session is busy, or that has an earlier request of the same session
// still waiting, is skipped (per-session order stays arrival order), and so
// is one whose device bytes do not fit next to the requests running.  Every
// session owns a HIP stream, so admitted requests of different sessions run
// concurrently on the device; a request's results are in the caller's
// buffers when its ticket completes.
#include "nckde_internal.h"

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

using nck::set_error;

namespace
{

struct Job
{
  nckde_request r;
  uint64_t ticket = 0;
  nckde_request_info info;
  bool done = false;
};

double
seconds_since (std::chrono::steady_clock::time_point t0)
{
  return std::chrono::duration<double> (std::chrono::steady_clock::now () - t0).count ();
}

} // namespace

struct nckde_server
{
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job *> pending;                  // arrival order
  std::unordered_map<uint64_t, Job *> jobs;   // submitted, not yet waited
  std::unordered_set<const nckde *> busy;     // sessions with a running request
  size_t budget = 0, in_flight = 0;
  unsigned running = 0;
  bool stop = false;
  uint64_t next_ticket = 1;
  int device = 0;
  std::chrono::steady_clock::time_point t0;
  std::vector<std::thread> workers;

  // first admissible pending request, or nullptr; mu held
  Job *
  pick ()
  {
    std::unordered_set<const nckde *> blocked;

    for (auto it = pending.begin (); it != pending.end (); ++it)
    {
      Job *j           = *it;
      const nckde *h   = j->r.h;

      if (busy.count (h) || blocked.count (h))
      {
        blocked.insert (h);
        continue;
      }

      if (running == 0 || in_flight + j->info.bytes <= budget)
      {
        pending.erase (it);
        return j;
      }

      blocked.insert (h);
    }

    return nullptr;
  }

  static int
  execute (const nckde_request &r)
  {
    unsigned d = 0;
    int st     = nckde_get_dim (r.h, &d);

    if (st != NCKDE_OK)
      return st;

    switch (r.op)
    {
      case NCKDE_REQ_EVAL_M2LNP:
        return nckde_eval_m2lnp (r.h, r.X, r.m, d, r.out, r.precision);
      case NCKDE_REQ_SAMPLE:
        return nckde_sample (r.h, r.u, r.z, r.chi2, r.m, r.out, r.idx);
      case NCKDE_REQ_PREPARE_INTERP:
        if (r.X != nullptr)
        {
          if ((st = nckde_reset (r.h)) != NCKDE_OK)
            return st;

          if ((st = nckde_add_obs_batch (r.h, r.X, r.m, d)) != NCKDE_OK)
            return st;
        }

        if ((st = nckde_prepare_interp (r.h, r.m2lnp, r.m)) != NCKDE_OK)
          return st;

        if (r.out != nullptr)
        {
          size_t nk = 0;

          if ((st = nckde_get_n_kernels (r.h, &nk)) != NCKDE_OK)
            return st;

          if ((st = nckde_get_weights (r.h, r.out)) != NCKDE_OK)
            return st;

          for (size_t i = nk; i < r.m; ++i)
            r.out[i] = std::numeric_limits<double>::quiet_NaN ();
        }

        return nck::trim_work (r.h);
      default:
        set_error ("nckde_server: unknown op %d", r.op);
        return NCKDE_EINVAL;
    }
  }

  void
  worker ()
  {
    (void) hipSetDevice (device);
    std::unique_lock<std::mutex> lk (mu);

    while (true)
    {
      Job *j = pick ();

      if (j == nullptr)
      {
        if (stop && pending.empty ())
          return;

        cv_work.wait (lk);
        continue;
      }

      busy.insert (j->r.h);
      in_flight += j->info.bytes;
      ++running;
      j->info.t_start = seconds_since (t0);
      lk.unlock ();

      const int st = execute (j->r);

      lk.lock ();
      j->info.t_end  = seconds_since (t0);
      j->info.status = st;

      if (st != NCKDE_OK)
        snprintf (j->info.message, sizeof (j->info.message), "%s", nckde_last_error ());

      busy.erase (j->r.h);
      in_flight -= j->info.bytes;
      --running;
      j->done = true;
      cv_done.notify_all ();
      cv_work.notify_all ();
    }
  }
};

extern "C" {

int
nckde_server_create (nckde_server **out, unsigned slots, size_t hbm_budget)
{
  if (out == nullptr || slots == 0 || slots > 64)
  {
    set_error ("nckde_server_create: slots must be in [1, 64]");
    return NCKDE_EINVAL;
  }

  *out = nullptr;

  nckde_server *s = new nckde_server ();

  if (hipGetDevice (&s->device) != hipSuccess)
  {
    set_error ("nckde_server_create: no HIP device available");
    delete s;
    return NCKDE_EHIP;
  }

  s->budget = hbm_budget;
  s->t0     = std::chrono::steady_clock::now ();

  for (unsigned i = 0; i < slots; ++i)
    s->workers.emplace_back ([s] { s->worker (); });

  *out = s;

  return NCKDE_OK;
}

int
nckde_server_destroy (nckde_server *s)
{
  if (s == nullptr)
    return NCKDE_OK;

  {
    std::lock_guard<std::mutex> lk (s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all ();

  for (auto &t : s->workers)
    t.join ();

  for (auto &kv : s->jobs)
    delete kv.second;

  delete s;

  return NCKDE_OK;
}

int
nckde_server_submit (nckde_server *s, const nckde_request *req, uint64_t *ticket)
{
  if (s == nullptr || req == nullptr || ticket == nullptr)
  {
    set_error ("nckde_server_submit: NULL argument");
    return NCKDE_EINVAL;
  }

  size_t bytes = 0;
  int st       = nckde_request_bytes (req, &bytes);

  if (st != NCKDE_OK)
    return st;

  const bool bad = (req->op == NCKDE_REQ_EVAL_M2LNP && (req->X == nullptr || req->out == nullptr)) ||
                   (req->op == NCKDE_REQ_SAMPLE && (req->u == nullptr || req->z == nullptr || req->out == nullptr)) ||
                   (req->op == NCKDE_REQ_PREPARE_INTERP && req->m2lnp == nullptr);

  if (bad)
  {
    set_error ("nckde_server_submit: missing buffer for op %d", req->op);
    return NCKDE_EINVAL;
  }

  Job *j = new Job ();

  j->r = *req;
  std::memset (&j->info, 0, sizeof (j->info));
  j->info.bytes = bytes;

  {
    std::lock_guard<std::mutex> lk (s->mu);

    if (s->stop)
    {
      delete j;
      set_error ("nckde_server_submit: the server is stopping");
      return NCKDE_ESTATE;
    }

    j->ticket          = s->next_ticket++;
    j->info.t_submit   = seconds_since (s->t0);
    s->jobs[j->ticket] = j;
    s->pending.push_back (j);
    *ticket            = j->ticket;
  }
  s->cv_work.notify_one ();

  return NCKDE_OK;
}

int
nckde_server_wait (nckde_server *s, uint64_t ticket, nckde_request_info *info)
{
  if (s == nullptr)
  {
    set_error ("nckde_server_wait: NULL server");
    return NCKDE_EINVAL;
  }

  std::unique_lock<std::mutex> lk (s->mu);
  auto it = s->jobs.find (ticket);

  if (it == s->jobs.end ())
  {
    set_error ("nckde_server_wait: unknown ticket %llu", (unsigned long long) ticket);
    return NCKDE_EINVAL;
  }

  Job *j = it->second;

  s->cv_done.wait (lk, [j] { return j->done; });
  s->jobs.erase (it);
  lk.unlock ();

  const int st = j->info.status;

  if (info)
    *info = j->info;

  if (st != NCKDE_OK)
    set_error ("%s", j->info.message);

  delete j;

  return st;
}

double
nckde_server_clock (nckde_server *s)
{
  return (s == nullptr) ? 0.0 : seconds_since (s->t0);
}

} // extern "C"
