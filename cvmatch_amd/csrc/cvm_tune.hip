// Measured launch-shape choices (the autotuner) and their cache file.
//
// Choices that change no output bit -- the direct form's waves per
// workgroup, the normalisation's strip width, the direct-vs-FFT crossover --
// are made per key (the device plus the shapes the choice depends on) by
// timing each candidate on the call's own data the first time the key comes
// up; the winner is kept for the process and written to a cache file, so
// later processes on the same device start from it.  Environment knobs force
// a choice:
//   CVMATCH_TUNE=0            no measuring: every key takes its default
//   CVMATCH_TUNE_FILE=path    the cache file (default
//                             $XDG_CACHE_HOME/cvmatch_hip/tune.txt, else
//                             ~/.cache/cvmatch_hip/tune.txt; "" = none)
//   CVMATCH_DIRECT_WAVES=1|2|4    waves per direct workgroup
//   CVMATCH_NORM_Q=2|4            result columns per lane of the scoring strips
//   CVMATCH_CROSSOVER=px      template side from which the mixed Fleet may
//                             route members to the FFT path (0: never)
//   CVMATCH_FLEET_FUSED=0|1   the direct Fleet's fused (1) or unfused launches
//   CVMATCH_FLEET_MERGE=0|1   fused member groups merged per (K step, square
//                             width) (1) or one launch per chunk count (0)
// The cache file holds one line per key: "key choice t0 t1 ...", the times
// (ms) of the candidates in the order they were tried; a key ends with the
// device and the library build it was measured with.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "cvm_internal.h"

namespace cvm {

namespace {

struct Decision {
  int choice = 0;
  std::vector<int> cands;
  std::vector<float> ms;
  std::string source;  // measured / cache / override / default
};

struct Tuner {
  std::mutex mu;
  bool loaded = false;
  std::string path, device;
  std::map<std::string, Decision> dec;

  static std::string cache_path() {
    const char* f = getenv("CVMATCH_TUNE_FILE");
    if (f) return f;
    const char* x = getenv("XDG_CACHE_HOME");
    std::string base;
    if (x && *x) {
      base = x;
    } else {
      const char* h = getenv("HOME");
      if (!h || !*h) return "";
      base = std::string(h) + "/.cache";
    }
    return base + "/cvmatch_hip/tune.txt";
  }

  void load() {
    if (loaded) return;
    loaded = true;
    int d = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&d) == hipSuccess && hipGetDeviceProperties(&p, d) == hipSuccess)
      device = std::string(p.gcnArchName) + "/" + std::to_string(p.multiProcessorCount);
    else
      device = "unknown";
    // choices measured by another build of the library (other kernels) are
    // not reused: the key carries this build's stamp
    device += "/" __DATE__ "/" __TIME__;
    for (char& c : device)
      if (c == ' ' || c == ':') c = '_';
    path = cache_path();
    if (path.empty()) return;
    FILE* fp = fopen(path.c_str(), "r");
    if (!fp) return;
    char line[1024];
    while (fgets(line, sizeof(line), fp)) {
      char key[512];
      int choice = 0, n = 0;
      if (sscanf(line, "%511s %d%n", key, &choice, &n) != 2) continue;
      Decision D;
      D.choice = choice;
      D.source = "cache";
      const char* q = line + n;
      float t;
      int k;
      while (sscanf(q, "%f%n", &t, &k) == 1) {
        D.ms.push_back(t);
        q += k;
      }
      dec[key] = D;
    }
    fclose(fp);
  }

  void save() {
    if (path.empty()) return;
    std::string dir = path.substr(0, path.find_last_of('/'));
    for (size_t i = 1; i <= dir.size(); i++)
      if (i == dir.size() || dir[i] == '/') mkdir(dir.substr(0, i).c_str(), 0755);
    std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    FILE* fp = fopen(tmp.c_str(), "w");
    if (!fp) return;
    for (auto& kv : dec) {
      if (kv.second.source != "measured" && kv.second.source != "cache") continue;
      fprintf(fp, "%s %d", kv.first.c_str(), kv.second.choice);
      for (float t : kv.second.ms) fprintf(fp, " %.5f", t);
      fprintf(fp, "\n");
    }
    fclose(fp);
    (void)rename(tmp.c_str(), path.c_str());
  }
};

Tuner& tuner() {
  static Tuner* t = new Tuner();
  return *t;
}

int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  if (!v || !*v) return dflt;
  return atoi(v);
}

}  // namespace

bool tuning_enabled() { return env_int("CVMATCH_TUNE", 1) != 0; }

// The lock is not held while candidates are measured: a measurement may run
// calls that consult the tuner themselves (the crossover's direct calls pick
// their waves).  Two threads measuring one key at once both store a choice;
// either is correct, as no choice changes an output bit.
int tune_choice(const std::string& what, const std::vector<int>& cands, const char* override_env,
                const std::function<float(int)>& time_of) {
  Tuner& T = tuner();
  std::string key;
  {
    std::lock_guard<std::mutex> lk(T.mu);
    T.load();
    key = what + "@" + T.device;
    const int forced = override_env ? env_int(override_env, -1) : -1;
    for (int c : cands)
      if (c == forced) {
        Decision D;
        D.choice = c;
        D.cands = cands;
        D.source = "override";
        T.dec[key] = D;
        return c;
      }
    auto it = T.dec.find(key);
    if (it != T.dec.end())
      for (int c : cands)
        if (c == it->second.choice) {
          it->second.cands = cands;
          return c;
        }
  }
  Decision D;
  D.cands = cands;
  D.choice = cands[0];
  D.source = "default";
  if (tuning_enabled() && cands.size() > 1 && time_of) {
    float best = -1.0f;
    for (int c : cands) {
      const float t = time_of(c);
      D.ms.push_back(t);
      if (t >= 0.0f && (best < 0.0f || t < best)) {
        best = t;
        D.choice = c;
      }
    }
    if (best >= 0.0f) D.source = "measured";
  }
  std::lock_guard<std::mutex> lk(T.mu);
  T.dec[key] = D;
  if (D.source == "measured") T.save();
  return D.choice;
}

int tune_value(const std::string& what, const char* override_env, int dflt,
               const std::function<int(std::vector<float>&)>& measure) {
  Tuner& T = tuner();
  std::string key;
  {
    std::lock_guard<std::mutex> lk(T.mu);
    T.load();
    key = what + "@" + T.device;
    const int forced = override_env ? env_int(override_env, -1) : -1;
    if (forced >= 0) {
      Decision D;
      D.choice = forced;
      D.source = "override";
      T.dec[key] = D;
      return forced;
    }
    auto it = T.dec.find(key);
    if (it != T.dec.end()) return it->second.choice;
  }
  Decision D;
  D.choice = dflt;
  D.source = "default";
  if (tuning_enabled() && measure) {
    D.choice = measure(D.ms);
    D.source = "measured";
  }
  std::lock_guard<std::mutex> lk(T.mu);
  T.dec[key] = D;
  if (D.source == "measured") T.save();
  return D.choice;
}

void tune_note(const std::string& what, int choice, const std::vector<float>& ms, const char* source) {
  Tuner& T = tuner();
  std::lock_guard<std::mutex> lk(T.mu);
  T.load();
  Decision D;
  D.choice = choice;
  D.ms = ms;
  D.source = source;
  T.dec[what + "@" + T.device] = D;
}

int tune_report(char* buf, int cap) {
  Tuner& T = tuner();
  std::lock_guard<std::mutex> lk(T.mu);
  std::string s = "{";
  bool first = true;
  for (auto& kv : T.dec) {
    char head[640];
    snprintf(head, sizeof(head), "%s\"%s\": {\"choice\": %d, \"source\": \"%s\", \"ms\": [", first ? "" : ", ",
             kv.first.c_str(), kv.second.choice, kv.second.source.c_str());
    s += head;
    for (size_t i = 0; i < kv.second.ms.size(); i++) {
      char t[32];
      snprintf(t, sizeof(t), "%s%.4f", i ? ", " : "", kv.second.ms[i]);
      s += t;
    }
    s += "], \"candidates\": [";
    for (size_t i = 0; i < kv.second.cands.size(); i++) s += (i ? ", " : "") + std::to_string(kv.second.cands[i]);
    s += "]}";
    first = false;
  }
  s += "}";
  if (buf && cap > 0) {
    snprintf(buf, (size_t)cap, "%s", s.c_str());
  }
  return (int)s.size();
}

// Average ms of one enqueue of `launch` on s: one untimed run, then `reps`
// timed between events (the stream is synchronised).  -1 when it fails.
float time_launches(hipStream_t s, const std::function<int()>& launch, int reps) {
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    (void)hipGetLastError();
    return -1.0f;
  }
  float ms = -1.0f;
  if (launch() == 0 && hipEventRecord(a, s) == hipSuccess) {
    bool ok = true;
    for (int r = 0; r < reps && ok; r++) ok = launch() == 0;
    if (ok && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
        hipEventElapsedTime(&ms, a, b) == hipSuccess)
      ms /= (float)reps;
    else
      ms = -1.0f;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipGetLastError();
  return ms;
}

}  // namespace cvm

// Self-test of the tuner's locking (host only, tests): a measured value
// whose measurement consults the tuner for another key, as the mixed form's
// crossover measurement does through its direct calls.  Returns the nested
// choice (2: the faster of the two fake candidates).
extern "C" int cvm_selftest_tuner(void) {
  using namespace cvm;
  int inner = -1;
  const int v = tune_value("selftest.outer", nullptr, 0, [&](std::vector<float>& ms) {
    inner = tune_choice("selftest.inner", {1, 2}, nullptr, [](int c) { return c == 2 ? 1.0f : 2.0f; });
    ms.push_back(0.0f);
    return 7;
  });
  return v == 7 ? inner : -1;
}
