// api_serve.cpp -- request server over nckde sessions (continuous batching).
//
// Worker threads take requests in arrival order, with backfill: a request
// whose session is busy, or that has an earlier request of the same session
// still waiting, is skipped (per-session order stays arrival order), and so
// is one whose device bytes do not fit next to the requests running.  Every
// session owns a HIP stream, so admitted requests of different sessions run
// concurrently on the device; a request's results are in the caller's
// buffers when its ticket completes.  A worker that takes a mixed eval also
// takes the admissible mixed evals of other sessions with the same fusion
// key and runs them as one fused launch per pass.
#include "nckde_internal.h"

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
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
  bool done     = false;
  size_t charge = 0;   // device bytes charged while it runs
};

// evals fused into one launch at most
constexpr size_t MAX_FUSED = 16;

// Refit buffers a worker keeps between its prepare_interp requests
struct WorkerState
{
  nck::WorkSet set;
  bool active = false;  // running a request (its set may be lent to a session)
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
  std::vector<std::unique_ptr<WorkerState>> wstate;

  // device bytes of the refit buffers held by idle workers; mu held
  size_t
  held_idle () const
  {
    size_t t = 0;

    for (const auto &w : wstate)
      if (!w->active)
        t += w->set.bytes ();

    return t;
  }

  // bytes charged for running j on worker wid: the request's own bytes plus,
  // for requests that do not use it, the refit set the worker holds
  size_t
  charge (const Job *j, unsigned wid) const
  {
    return j->info.bytes + ((j->r.op == NCKDE_REQ_PREPARE_INTERP) ? 0 : wstate[wid]->set.bytes ());
  }

  // first admissible pending request for worker wid, or nullptr; mu held.
  // When only the idle workers' refit sets keep a request out, they are
  // freed (the one place where serving frees device memory).
  Job *
  pick (unsigned wid)
  {
    for (int pass = 0; pass < 2; ++pass)
    {
      std::unordered_set<const nckde *> blocked;
      const size_t held = held_idle () - wstate[wid]->set.bytes ();
      bool held_blocks  = false;

      for (auto it = pending.begin (); it != pending.end (); ++it)
      {
        Job *j         = *it;
        const nckde *h = j->r.h;

        if (busy.count (h) || blocked.count (h))
        {
          blocked.insert (h);
          continue;
        }

        const size_t c = charge (j, wid);

        if (running == 0 || in_flight + held + c <= budget)
        {
          pending.erase (it);
          return j;
        }

        if (in_flight + c <= budget)
          held_blocks = true;

        blocked.insert (h);
      }

      if (!held_blocks || pass == 1)
        break;

      for (auto &w : wstate)
        if (!w->active)
          w->set.release ();
    }

    return nullptr;
  }

  // the fusion key of a waiting request (-1: runs alone)
  static int
  fuse_key (const Job *j)
  {
    if (j->r.op != NCKDE_REQ_EVAL_M2LNP || j->r.precision != NCKDE_PREC_MIXED)
      return -1;

    return nck::mixed_fuse_key (j->r.h);
  }

  // moves the waiting mixed evals that can share j's launch into batch
  // (after j): their sessions idle and not behind an earlier waiting
  // request, one per session, within the budget; mu held
  void
  gather_fused (Job *j, std::vector<Job *> &batch)
  {
    batch.assign (1, j);

    const int key = fuse_key (j);

    if (key < 0)
      return;

    std::unordered_set<const nckde *> blocked = {j->r.h};
    size_t extra                              = 0;

    for (auto it = pending.begin (); it != pending.end () && batch.size () < MAX_FUSED;)
    {
      Job *c         = *it;
      const nckde *h = c->r.h;

      if (busy.count (h) || blocked.count (h) || fuse_key (c) != key ||
          in_flight + j->charge + extra + c->info.bytes > budget)
      {
        blocked.insert (h);
        ++it;
        continue;
      }

      c->charge = c->info.bytes;
      extra += c->charge;
      blocked.insert (h);
      batch.push_back (c);
      it = pending.erase (it);
    }
  }

  static int
  execute_fused (const std::vector<Job *> &batch)
  {
    const size_t nb = batch.size ();
    std::vector<nckde *> hs (nb);
    std::vector<const double *> X (nb);
    std::vector<size_t> m (nb);
    std::vector<double *> out (nb);

    for (size_t k = 0; k < nb; ++k)
    {
      hs[k]  = batch[k]->r.h;
      X[k]   = batch[k]->r.X;
      m[k]   = batch[k]->r.m;
      out[k] = batch[k]->r.out;
    }

    return nckde_eval_m2lnp_multi (hs.data (), X.data (), m.data (), nullptr, out.data (), nb, NCKDE_PREC_MIXED);
  }

  static int
  execute (const nckde_request &r, nck::WorkSet &ws)
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

        // the refit runs on the worker's buffers, which go back to the
        // worker afterwards (also when the refit fails)
        if ((st = nck::swap_work (r.h, ws)) != NCKDE_OK)
          return st;

        st = nckde_prepare_interp (r.h, r.m2lnp, r.m);

        {
          const int st2 = nck::swap_work (r.h, ws);

          if (st == NCKDE_OK)
            st = st2;
        }

        if (st != NCKDE_OK)
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

        return NCKDE_OK;
      default:
        set_error ("nckde_server: unknown op %d", r.op);
        return NCKDE_EINVAL;
    }
  }

  void
  worker (unsigned wid)
  {
    (void) hipSetDevice (device);
    std::unique_lock<std::mutex> lk (mu);
    WorkerState &me = *wstate[wid];
    std::vector<Job *> batch;

    while (true)
    {
      Job *j = pick (wid);

      if (j == nullptr)
      {
        if (stop && pending.empty ())
          return;

        cv_work.wait (lk);
        continue;
      }

      j->charge = charge (j, wid);
      gather_fused (j, batch);

      const double t_start = seconds_since (t0);

      for (Job *b : batch)
      {
        busy.insert (b->r.h);
        in_flight += b->charge;
        ++running;
        b->info.t_start = t_start;
        b->info.batch   = (unsigned) batch.size ();
      }

      me.active = true;
      lk.unlock ();

      const int st = (batch.size () > 1) ? execute_fused (batch) : execute (j->r, me.set);
      char msg[sizeof (j->info.message)] = "";

      if (st != NCKDE_OK)
        snprintf (msg, sizeof (msg), "%s", nckde_last_error ());

      lk.lock ();
      me.active = false;

      const double t_end = seconds_since (t0);

      for (Job *b : batch)
      {
        b->info.t_end  = t_end;
        b->info.status = st;
        memcpy (b->info.message, msg, sizeof (msg));
        busy.erase (b->r.h);
        in_flight -= b->charge;
        --running;
        b->done = true;
      }

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
    s->wstate.emplace_back (new WorkerState ());

  for (unsigned i = 0; i < slots; ++i)
    s->workers.emplace_back ([s, i] { s->worker (i); });

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

  for (auto &w : s->wstate)
    w->set.release ();

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

  // the ticket leaves the table before the wait, so a second waiter on it
  // gets "unknown ticket" instead of a job another thread frees
  Job *j = it->second;

  s->jobs.erase (it);
  s->cv_done.wait (lk, [j] { return j->done; });
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
