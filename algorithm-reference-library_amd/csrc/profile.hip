// Opt-in per-kernel timing with HIP events recorded on the launching stream.
//
// When enabled, every library kernel launch is bracketed by two events on its stream;
// arlmi_profile_collect() waits for them and accumulates the elapsed time per kernel name.
#include "arlmi_internal.h"

#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

namespace arlmi {

namespace {
struct Pending {
  const char* name;
  hipEvent_t a, b;
};
struct Totals {
  double ms = 0.0;
  long count = 0;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_pool;
std::map<std::string, Totals> g_tot;
std::set<std::string> g_only;  // kernels to time; empty = all

hipEvent_t take() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

bool profile_on() { return g_on; }

void* profile_begin(const char* name, hipStream_t s) {
  if (!g_on) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_only.empty() && g_only.find(name) == g_only.end()) return nullptr;
  Pending* p = new Pending{name, take(), take()};
  if (p->a) (void)hipEventRecord(p->a, s);
  return p;
}

void profile_end(void* token, hipStream_t s) {
  if (!token) return;
  Pending* p = static_cast<Pending*>(token);
  std::lock_guard<std::mutex> lk(g_mu);
  if (p->b) (void)hipEventRecord(p->b, s);
  g_pending.push_back(*p);
  delete p;
}

}  // namespace arlmi

using namespace arlmi;

extern "C" {

int arlmi_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_on = on != 0;
  return ARLMI_OK;
}

// Restrict the timing to the kernels named in a comma-separated list (NULL or "" = all).  Each
// timed launch puts two event markers on its stream, which cost a few microseconds of stream
// time each; timing only the long kernels keeps that out of short launch sequences.
int arlmi_profile_only(const char* names) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_only.clear();
  if (!names) return ARLMI_OK;
  std::string cur;
  for (const char* c = names;; ++c) {
    if (*c == ',' || *c == 0) {
      if (!cur.empty()) g_only.insert(cur);
      cur.clear();
      if (*c == 0) break;
    } else if (*c != ' ') {
      cur += *c;
    }
  }
  return ARLMI_OK;
}

// Wait for the recorded events and fold them into the per-name totals.
int arlmi_profile_collect(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& p : g_pending) {
    float ms = 0.0f;
    if (p.a && p.b && hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      Totals& t = g_tot[p.name];
      t.ms += ms;
      t.count += 1;
    }
    if (p.a) g_pool.push_back(p.a);
    if (p.b) g_pool.push_back(p.b);
  }
  g_pending.clear();
  return ARLMI_OK;
}

int arlmi_profile_reset(void) {
  arlmi_profile_collect();
  std::lock_guard<std::mutex> lk(g_mu);
  g_tot.clear();
  return ARLMI_OK;
}

// Copies "name ms count\n" lines into buf (NUL-terminated); returns the bytes needed.
int arlmi_profile_report(char* buf, int cap) {
  std::lock_guard<std::mutex> lk(g_mu);
  std::string s;
  for (auto& kv : g_tot) {
    char line[256];
    snprintf(line, sizeof(line), "%s %.6f %ld\n", kv.first.c_str(), kv.second.ms, kv.second.count);
    s += line;
  }
  if (buf && cap > 0) {
    const int n = (int)s.size() < cap - 1 ? (int)s.size() : cap - 1;
    memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return (int)s.size() + 1;
}

}  // extern "C"
