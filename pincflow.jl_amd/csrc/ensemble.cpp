// Ensembles on one GPU (src/ensemble.jl:20-74 runs each member in an MPI
// communicator of its own).  Each member is a state stepped by a host thread
// of its own; attached to a pcf_ensemble, its BiCGSTAB solves run in the
// ensemble's batch.  A service thread owns the batch: it admits a member's
// solve at the next iteration boundary (the solve's vectors and scalars are
// set up on the member's stream by bicgstab_core, and the service's stream
// waits for them), launches one iteration of every solve in progress as one
// launch per kernel whose grid covers these solves (a list of their slots in
// the member table of each solve kind), and releases a solve as soon as its
// convergence flag (or its own poisson_iterations) is seen, one iteration
// later.  Members step
// independently: a member whose solve ends goes on with its explicit work
// while the others' solves continue, and the batch is never held for the
// slowest member.  A member's results are those of its own solve (the same
// blocks, the same partial sums, the same bits); only the launches are
// shared.  Solves that the table does not cover (the multigrid or
// learned-stage preconditioners, profiled runs) run on their own.
#include <algorithm>

#include "pcf_internal.hpp"

namespace pcf {

namespace {

// Records the service's first error and fails every solve in the batch and
// every waiting one (lock held).
void fail_all(pcf_ensemble* e, int rc) {
  if (!e->err) e->err = rc;
  for (auto& g : e->groups) {
    for (pcf_state* m : g->active) {
      m->ens_active = false;
      m->ens_rc = rc;
    }
    g->active.clear();
  }
  for (pcf_state* m : e->pending) m->ens_rc = rc;
  e->pending.clear();
  e->done_cv.notify_all();
}

bool any_active(const pcf_ensemble* e) {
  for (auto& g : e->groups)
    if (!g->active.empty()) return true;
  return false;
}

void service_loop(pcf_ensemble* e) {
  std::unique_lock<std::mutex> lk(e->mu);
  if (hipSetDevice(e->device) != hipSuccess) e->err = PCF_E_HIP;
  long long pass = 0;
  for (;;) {
    while (!e->stop && e->pending.empty() && !any_active(e)) e->wake.wait(lk);
    if (e->pending.empty() && !any_active(e)) break;  // stop
    if (e->err) {
      fail_all(e, e->err);
      continue;
    }
    int rc = PCF_OK;
    // tables of groups whose members changed (nothing in flight after the
    // upload's synchronisation); slots may move, so the previous pass's
    // flags are not read in this one
    bool rebuilt = false;
    for (auto& g : e->groups)
      if (g->stale && !g->slots.empty()) {
        if ((rc = ensemble_upload(e, *g))) break;
        rebuilt = true;
      }
    // admission: the batch's stream waits for the member's set-up
    for (size_t q = 0; q < e->pending.size() && rc == PCF_OK; ++q) {
      pcf_state* m = e->pending[q];
      if (hipStreamWaitEvent(e->stream, m->ens_ready, 0) != hipSuccess) rc = PCF_E_HIP;
      m->ens_active = true;
      m->ens_since = pass + 1;
      e->groups[m->ens_group]->active.push_back(m);
    }
    e->pending.clear();
    if (rc) {
      fail_all(e, rc);
      continue;
    }
    ++pass;
    // one iteration of every solve in the batch (the lock is held: member
    // threads only queue solves or attach meanwhile), then the done flags
    int32_t* fl = e->flags_host + (pass & 1) * kEnsGroups * kEnsMax;
    for (size_t gi = 0; gi < e->groups.size() && rc == PCF_OK; ++gi)
      rc = ensemble_iterate(e, *e->groups[gi], (int)(pass & 1), e->flags_dev + gi * kEnsMax);
    if (rc == PCF_OK &&
        (hipMemcpyAsync(fl, e->flags_dev, e->groups.size() * kEnsMax * sizeof(int32_t), hipMemcpyDeviceToHost,
                        e->stream) != hipSuccess ||
         hipEventRecord(e->ev[pass & 1], e->stream) != hipSuccess))
      rc = PCF_E_HIP;
    if (rc) {
      fail_all(e, rc);
      continue;
    }
    e->iterations += 1;
    if (pass < 2) continue;
    // the previous iteration's flags, for the solves it ran
    lk.unlock();
    const bool waited = hipEventSynchronize(e->ev[(pass - 1) & 1]) == hipSuccess;
    lk.lock();
    if (!waited) {
      fail_all(e, PCF_E_HIP);
      continue;
    }
    if (rebuilt) continue;
    const int32_t* seen = e->flags_host + ((pass - 1) & 1) * kEnsGroups * kEnsMax;
    bool released = false;
    for (auto& g : e->groups) {
      for (size_t q = 0; q < g->active.size();) {
        pcf_state* m = g->active[q];
        if (m->ens_since > pass - 1 || seen[m->ens_group * kEnsMax + m->ens_slot] == 0) {
          ++q;
          continue;
        }
        // converged or stopped (its blocks in this pass's iteration returned
        // at once): out of the next iterations, its final scalars to the
        // host, and its stream may go on after this point
        if (hipMemcpyAsync(m->sc_host, m->sc, sizeof(SolverScalars), hipMemcpyDeviceToHost, e->stream) !=
                hipSuccess ||
            hipEventRecord(m->ens_done, e->stream) != hipSuccess) {
          rc = PCF_E_HIP;
          break;
        }
        m->ens_active = false;
        m->ens_rc = PCF_OK;
        e->member_solves += 1;
        g->active.erase(g->active.begin() + q);
        released = true;
      }
      if (rc) break;
    }
    if (rc) fail_all(e, rc);
    else if (released) e->done_cv.notify_all();
  }
}

}  // namespace

// The member's side: hand the solve (set up on this state's stream) to the
// service and wait for its release.
int ensemble_solve(pcf_state* st) {
  pcf_ensemble* e = st->ens;
  PCF_HIP(hipEventRecord(st->ens_ready, st->stream));
  std::unique_lock<std::mutex> lk(e->mu);
  if (e->err) return e->err;
  st->ens_rc = -1;
  e->pending.push_back(st);
  e->wake.notify_one();
  e->done_cv.wait(lk, [&] { return st->ens_rc >= 0; });
  const int rc = st->ens_rc;
  lk.unlock();
  if (rc) return rc;
  // the batch's last work on this state (and the copy of its scalars)
  PCF_HIP(hipEventSynchronize(st->ens_done));
  return PCF_OK;
}

}  // namespace pcf

using namespace pcf;

namespace {

// The ensemble's device resources and service thread, made by the first
// attach (on the device of the first member; lock held).
int ensemble_start(pcf_ensemble* e, const pcf_state* st) {
  if (e->stream) return PCF_OK;
  e->device = st->device;
  PCF_HIP(hipSetDevice(e->device));
  PCF_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  PCF_HIP(hipMalloc((void**)&e->flags_dev, kEnsGroups * kEnsMax * sizeof(int32_t)));
  PCF_HIP(hipMemset(e->flags_dev, 0, kEnsGroups * kEnsMax * sizeof(int32_t)));
  PCF_HIP(hipHostMalloc((void**)&e->flags_host, 2 * kEnsGroups * kEnsMax * sizeof(int32_t), hipHostMallocDefault));
  for (int q = 0; q < 2; ++q) PCF_HIP(hipEventCreateWithFlags(&e->ev[q], hipEventDisableTiming));
  e->service = std::thread(service_loop, e);
  return PCF_OK;
}

}  // namespace

extern "C" {

int pcf_ensemble_create(pcf_ensemble** out) {
  if (!out) return PCF_E_ARG;
  *out = new pcf_ensemble();
  return PCF_OK;
}

int pcf_ensemble_destroy(pcf_ensemble* e) {
  if (!e) return PCF_OK;
  {
    std::unique_lock<std::mutex> lk(e->mu);
    if (e->attached > 0) {
      set_error("pcf_ensemble_destroy: members are still attached");
      return PCF_E_ARG;
    }
    e->stop = true;
    e->wake.notify_all();
  }
  if (e->service.joinable()) e->service.join();
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (auto& g : e->groups) {
    if (g->table) (void)hipFree(g->table);
    if (g->list) (void)hipFree(g->list);
    if (g->list_host) (void)hipHostFree(g->list_host);
  }
  for (int q = 0; q < 2; ++q)
    if (e->ev[q]) (void)hipEventDestroy(e->ev[q]);
  if (e->flags_dev) (void)hipFree(e->flags_dev);
  if (e->flags_host) (void)hipHostFree(e->flags_host);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return PCF_OK;
}

int pcf_ensemble_attach(pcf_ensemble* e, pcf_state* st) {
  if (!e || !st) return PCF_E_ARG;
  if (st->ens) {
    set_error("pcf_ensemble_attach: the state is attached to an ensemble already");
    return PCF_E_ARG;
  }
  for (hipEvent_t* ev : {&st->ens_ready, &st->ens_done})
    if (!*ev && hipEventCreateWithFlags(ev, hipEventDisableTiming) != hipSuccess) {
      set_error("pcf_ensemble_attach: hipEventCreateWithFlags failed");
      return PCF_E_HIP;
    }
  std::unique_lock<std::mutex> lk(e->mu);
  if (e->attached >= kEnsMax) {
    set_error("pcf_ensemble_attach: at most " + std::to_string(kEnsMax) + " members");
    return PCF_E_ARG;
  }
  if (e->stream && st->device != e->device) {
    set_error("pcf_ensemble_attach: the members of an ensemble share one device");
    return PCF_E_ARG;
  }
  const int key = ensemble_key(st);
  int gi = -1;
  if (key >= 0) {
    for (size_t q = 0; q < e->groups.size(); ++q)
      if (e->groups[q]->key == key) gi = (int)q;
    if (gi < 0 && (int)e->groups.size() >= kEnsGroups) {
      set_error("pcf_ensemble_attach: at most " + std::to_string(kEnsGroups) + " solve kinds in one ensemble");
      return PCF_E_ARG;
    }
  }
  int rc = ensemble_start(e, st);
  if (rc) return rc;
  if (key >= 0) {
    if (gi < 0) {
      e->groups.emplace_back(new pcf_ensemble::Group());
      gi = (int)e->groups.size() - 1;
      e->groups[gi]->key = key;
    }
    pcf_ensemble::Group& g = *e->groups[gi];
    st->ens_slot = (int)g.slots.size();
    g.slots.push_back(st);
    g.stale = true;
  }
  st->ens_group = gi;
  st->ens_active = false;
  st->ens = e;
  e->attached += 1;
  return PCF_OK;
}

int pcf_ensemble_detach(pcf_state* st) {
  if (!st) return PCF_E_ARG;
  pcf_ensemble* e = st->ens;
  if (!e) return PCF_OK;
  std::unique_lock<std::mutex> lk(e->mu);
  // a member's thread detaches between its calls, so its solve is not in
  // the batch
  if (st->ens_group >= 0) {
    pcf_ensemble::Group& g = *e->groups[st->ens_group];
    g.slots.erase(std::find(g.slots.begin(), g.slots.end(), st));
    g.stale = true;
  }
  st->ens_group = st->ens_slot = -1;
  st->ens = nullptr;
  e->attached -= 1;
  return PCF_OK;
}

int pcf_ensemble_stats(pcf_ensemble* e, int64_t out[4]) {
  if (!e || !out) return PCF_E_ARG;
  std::unique_lock<std::mutex> lk(e->mu);
  out[0] = e->iterations;
  out[1] = e->member_solves;
  out[2] = e->solves_alone;
  out[3] = e->table_builds;
  return PCF_OK;
}

}  // extern "C"
