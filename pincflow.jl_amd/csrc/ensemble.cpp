// Ensembles on one GPU (src/ensemble.jl:20-74 runs each member in an MPI
// communicator of its own).  Each member is a state stepped by a host thread
// of its own; attached to a pcf_ensemble, the BiCGSTAB iterations of its
// solves run in batches with the other members': a member whose solve is
// ready waits until every attached member waits, and the last to arrive
// runs one launch per kernel of the iteration over all waiting members
// (ensemble_batch), replayed from a hipGraph, until each has converged or
// reached its poisson_iterations.  A member's results are those of its own
// solve (the same blocks, the same partial sums, the same bits); only the
// launches are shared.  Solves that the batch does not cover (the multigrid
// or learned-stage preconditioners, profiled runs) step out of the count
// while they run on their own.
#include <algorithm>

#include "pcf_internal.hpp"

namespace pcf {

namespace {

// Runs the batches of everything waiting (lock held), sets each member's
// result and wakes the waiting threads.
void run_batches(pcf_ensemble* e) {
  std::vector<pcf_state*> waiting;
  waiting.swap(e->waiting);
  std::vector<int> keys;
  for (pcf_state* m : waiting) {
    const int k = ensemble_key(m, m->ens_mixed);
    if (std::find(keys.begin(), keys.end(), k) == keys.end()) keys.push_back(k);
  }
  for (int k : keys) {
    std::vector<pcf_state*> mem;
    for (pcf_state* m : waiting)
      if (ensemble_key(m, m->ens_mixed) == k) mem.push_back(m);
    // one order whatever the order of arrival, so that the table (and its
    // captured iteration) is rebuilt only when the members change
    std::sort(mem.begin(), mem.end());
    pcf_ensemble::Group* grp = nullptr;
    for (auto& g : e->groups)
      if (g->key == k) grp = g.get();
    if (!grp) {
      e->groups.emplace_back(new pcf_ensemble::Group());
      grp = e->groups.back().get();
      grp->key = k;
    }
    int rc = hipSetDevice(e->device) == hipSuccess ? ensemble_batch(e, *grp, mem) : PCF_E_HIP;
    for (pcf_state* m : mem) m->ens_rc = rc;
  }
  e->cv.notify_all();
}

}  // namespace

// The member's side: wait for (or run) the batch that takes this solve.
int ensemble_solve(pcf_state* st, bool mixed) {
  pcf_ensemble* e = st->ens;
  PCF_HIP(hipEventRecord(st->ens_ready, st->stream));
  std::unique_lock<std::mutex> lk(e->mu);
  st->ens_mixed = mixed;
  st->ens_rc = -1;
  e->waiting.push_back(st);
  while (st->ens_rc < 0) {
    if ((int)e->waiting.size() >= e->attached) run_batches(e);
    else e->cv.wait(lk);
  }
  return st->ens_rc;
}

// A solve outside the batch leaves the count while it runs.
void ensemble_step_out(pcf_state* st, bool out) {
  pcf_ensemble* e = st->ens;
  std::unique_lock<std::mutex> lk(e->mu);
  if (out) {
    e->attached -= 1;
    e->stepped_out += 1;
    if (!e->waiting.empty() && (int)e->waiting.size() >= e->attached) run_batches(e);
  } else {
    e->attached += 1;
  }
}

}  // namespace pcf

using namespace pcf;

extern "C" {

// The ensemble's device resources, made by the first attach (on the
// device of the first member; lock held).
static int ensemble_device_setup(pcf_ensemble* e, const pcf_state* st) {
  if (e->stream) return PCF_OK;
  e->device = st->device;
  PCF_HIP(hipSetDevice(e->device));
  PCF_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  PCF_HIP(hipMalloc((void**)&e->flags_dev, kEnsMax * sizeof(int32_t)));
  PCF_HIP(hipHostMalloc((void**)&e->flags_host, 2 * kEnsMax * sizeof(int32_t), hipHostMallocDefault));
  for (int q = 0; q < 2; ++q) PCF_HIP(hipEventCreateWithFlags(&e->ev[q], hipEventDisableTiming));
  return PCF_OK;
}

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
  }
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (auto& g : e->groups) {
    if (g->graph) (void)hipGraphExecDestroy(g->graph);
    if (g->table) (void)hipFree(g->table);
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
  if (!st->ens_ready && hipEventCreateWithFlags(&st->ens_ready, hipEventDisableTiming) != hipSuccess) {
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
  int rc = ensemble_device_setup(e, st);
  if (rc) return rc;
  e->attached += 1;
  st->ens = e;
  return PCF_OK;
}

int pcf_ensemble_detach(pcf_state* st) {
  if (!st) return PCF_E_ARG;
  pcf_ensemble* e = st->ens;
  if (!e) return PCF_OK;
  std::unique_lock<std::mutex> lk(e->mu);
  e->attached -= 1;
  st->ens = nullptr;
  // the members waiting for this one's solve need not wait any longer
  if (!e->waiting.empty() && (int)e->waiting.size() >= e->attached) run_batches(e);
  return PCF_OK;
}

int pcf_ensemble_stats(pcf_ensemble* e, int64_t out[4]) {
  if (!e || !out) return PCF_E_ARG;
  std::unique_lock<std::mutex> lk(e->mu);
  out[0] = e->batches;
  out[1] = e->iterations;
  out[2] = e->member_solves;
  out[3] = e->stepped_out;
  return PCF_OK;
}

int pcf_ensemble_set_graphs(pcf_ensemble* e, int32_t on) {
  if (!e) return PCF_E_ARG;
  std::unique_lock<std::mutex> lk(e->mu);
  e->graphs = on != 0;
  return PCF_OK;
}

}  // extern "C"
