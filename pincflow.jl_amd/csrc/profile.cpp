// Kernel timing with HIP events recorded on the state's stream around the
// launches of a kernel group.  Pending event pairs are resolved whenever the
// stream is idle (end of pcf_step, pcf_profile_query) and their events reused.
#include "pcf_internal.hpp"

namespace pcf {

namespace {
hipEvent_t take(pcf_state* st) {
  if (!st->prof_pool.empty()) {
    hipEvent_t e = st->prof_pool.back();
    st->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

void prof_begin(pcf_state* st, int group, hipEvent_t* a) {
  *a = nullptr;
  if (!st->prof_on || !((st->prof_mask >> group) & 1u)) return;
  *a = take(st);
  if (*a) (void)hipEventRecord(*a, st->stream);
}

void prof_end(pcf_state* st, int group, hipEvent_t a) {
  if (!st->prof_on || !a) return;
  hipEvent_t b = take(st);
  if (!b) return;
  (void)hipEventRecord(b, st->stream);
  st->prof_pending.push_back({group, a, b});
  if (st->prof_pending.size() > 4096) prof_flush(st);
}

void prof_flush(pcf_state* st) {
  for (auto& r : st->prof_pending) {
    (void)hipEventSynchronize(r.b);
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      st->prof_ms[r.group] += ms;
      st->prof_n[r.group] += 1;
    }
    st->prof_pool.push_back(r.a);
    st->prof_pool.push_back(r.b);
  }
  st->prof_pending.clear();
}

void prof_release(pcf_state* st) {
  prof_flush(st);
  for (hipEvent_t e : st->prof_pool) (void)hipEventDestroy(e);
  st->prof_pool.clear();
}

}  // namespace pcf

extern "C" {

int pcf_profile_groups(pcf_state* st, uint32_t mask) {
  if (!st) return PCF_E_ARG;
  pcf::prof_flush(st);
  mask &= (1u << PCF_K_COUNT) - 1u;
  if (mask) {
    for (int g = 0; g < PCF_K_COUNT; ++g) {
      st->prof_ms[g] = 0.0;
      st->prof_n[g] = 0;
    }
  }
  st->prof_on = mask != 0;
  st->prof_mask = mask;
  return PCF_OK;
}

int pcf_profile_enable(pcf_state* st, int on) { return pcf_profile_groups(st, on ? ~0u : 0u); }

int pcf_profile_query(pcf_state* st, int group, int64_t* launches, double* total_ms) {
  if (!st || group < 0 || group >= PCF_K_COUNT || !launches || !total_ms) return PCF_E_ARG;
  pcf::prof_flush(st);
  *launches = st->prof_n[group];
  *total_ms = st->prof_ms[group];
  return PCF_OK;
}

}  // extern "C"
