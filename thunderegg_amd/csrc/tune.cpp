// The autotuner's table (tune.h) and its C entry points.
#include "tune.h"

#include <dlfcn.h>

#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

#include "kernels/dft.h"
#include "te_internal.h"

namespace te {
namespace {

constexpr DftGenericVariant kGeneric[] = { { 1, 256 }, { 1, 128 }, { 1, 64 }, { 2, 128 },
                                           { 4, 64 },  { 2, 64 },  { 8, 32 } };
constexpr int kGenericCount = sizeof(kGeneric) / sizeof(kGeneric[0]);

using Key = std::array<int, 4>;  // family, dim, n, dtype

struct Table {
  std::mutex m;
  std::map<Key, int> v;
  std::string path;  // the file loaded ("" for none)
  bool loaded = false;
};

Table& table()
{
  static Table t;
  return t;
}

struct Forced {
  bool on = false;
  Key k{};
  int variant = 0;
};
thread_local Forced forced;

bool valid(int family, int variant)
{
  return family >= 0 && family < TUNE_FAMILIES && variant >= 0 && variant < tune_num_variants(family);
}

// parse "family dim n dtype variant" lines ('#' starts a comment); false on a bad line
bool parse(std::istream& in, std::map<Key, int>& out, std::string& err)
{
  std::string line;
  int ln = 0;
  while (std::getline(in, line)) {
    ln++;
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream ls(line);
    int f, d, n, dt, v;
    if (!(ls >> f)) continue;
    if (!(ls >> d >> n >> dt >> v) || !valid(f, v) || (d != 2 && d != 3) || n < 1 ||
        (dt != TE_F32 && dt != TE_F64)) {
      err = "line " + std::to_string(ln) + ": expected 'family dim n dtype variant' with a valid variant";
      return false;
    }
    out[Key{ f, d, n, dt }] = v;
  }
  return true;
}

int load_file(const std::string& path)
{
  std::ifstream in(path);
  if (!in) {
    set_error("te_tune_load: cannot open " + path);
    return TE_ERR_ARG;
  }
  std::map<Key, int> v;
  std::string err;
  if (!parse(in, v, err)) {
    set_error("te_tune_load: " + path + ": " + err);
    return TE_ERR_ARG;
  }
  Table& t = table();
  std::lock_guard<std::mutex> lk(t.m);
  t.v.swap(v);
  t.path = path;
  t.loaded = true;
  return TE_OK;
}

// the default table: TE_TUNE_TABLE, or tune_table.txt beside this library (a missing file
// leaves the hand defaults)
void load_default()
{
  static std::once_flag once;
  std::call_once(once, [] {
    const char* env = std::getenv("TE_TUNE_TABLE");
    std::string path;
    if (env && env[0]) {
      if (std::strcmp(env, "none") == 0) return;
      path = env;
    } else {
      Dl_info info;
      if (!dladdr((void*)&load_default, &info) || !info.dli_fname) return;
      std::string so = info.dli_fname;
      const size_t sl = so.rfind('/');
      path = (sl == std::string::npos ? std::string(".") : so.substr(0, sl)) + "/tune_table.txt";
      std::ifstream probe(path);
      if (!probe) return;
    }
    Table& t = table();
    {
      std::lock_guard<std::mutex> lk(t.m);
      if (t.loaded) return;  // te_tune_load ran first
    }
    if (load_file(path) != TE_OK) std::fprintf(stderr, "thunderegg_amd: ignoring tune table: %s\n", te_last_error());
  });
}

}  // namespace

int tune_num_variants(int family)
{
  if (family == TUNE_DFT_GENERIC) return kGenericCount;
  if (family == TUNE_DFT_MFMA_2D) return MFMA_GRID_VARIANTS;
  if (family == TUNE_C3_SWEEP) return kC3Variants;
  return 0;
}

DftGenericVariant dft_generic_variant(int v) { return kGeneric[v >= 0 && v < kGenericCount ? v : 0]; }

void tune_force(int family, int dim, int n, int dtype, int variant)
{
  forced.on = true;
  forced.k = Key{ family, dim, n, dtype };
  forced.variant = variant;
}

void tune_unforce() { forced.on = false; }

int tune_variant(int family, int dim, int n, int dtype)
{
  const Key k{ family, dim, n, dtype };
  if (forced.on && forced.k == k) return forced.variant;
  load_default();
  Table& t = table();
  std::lock_guard<std::mutex> lk(t.m);
  auto it = t.v.find(k);
  return it == t.v.end() ? 0 : it->second;
}

}  // namespace te

using namespace te;

extern "C" {

int te_tune_num_variants(int family) { return tune_num_variants(family); }

int te_tune_load(const char* path)
{
  if (!path) {
    set_error("te_tune_load: null path");
    return TE_ERR_ARG;
  }
  return load_file(path);
}

int te_tune_clear(void)
{
  Table& t = table();
  std::lock_guard<std::mutex> lk(t.m);
  t.v.clear();
  t.path.clear();
  t.loaded = true;
  return TE_OK;
}

int te_tune_set(int family, int dim, int n, int dtype, int variant)
{
  if (!valid(family, variant)) {
    set_error("te_tune_set: bad family or variant");
    return TE_ERR_ARG;
  }
  load_default();
  Table& t = table();
  std::lock_guard<std::mutex> lk(t.m);
  t.v[Key{ family, dim, n, dtype }] = variant;
  return TE_OK;
}

int te_tune_get(int family, int dim, int n, int dtype) { return tune_variant(family, dim, n, dtype); }

int te_tune_table_path(char* buf, size_t cap)
{
  load_default();
  Table& t = table();
  std::lock_guard<std::mutex> lk(t.m);
  if (!buf || cap == 0) return (int)t.path.size();
  std::snprintf(buf, cap, "%s", t.path.c_str());
  return (int)t.path.size();
}

int te_tune_time_dft(const te_dft_plan* pl, int dtype, const void* f, const void* u, void* out, int variant, int reps,
                     double* ms)
{
  if (!pl || !f || !u || !out || !ms || reps < 1 || (dtype != TE_F32 && dtype != TE_F64)) {
    set_error("te_tune_time_dft: bad argument");
    return TE_ERR_ARG;
  }
  const te_domain* d = pl->g->d;
  const int family = (pl->use_mfma && dft_mfma_supported(pl)) ? TUNE_DFT_MFMA_2D : TUNE_DFT_GENERIC;
  if (!valid(family, variant)) {
    set_error("te_tune_time_dft: bad variant");
    return TE_ERR_ARG;
  }
  hipStream_t s;
  hipEvent_t e0, e1;
  TE_HIP(hipStreamCreate(&s));
  TE_HIP(hipEventCreate(&e0));
  TE_HIP(hipEventCreate(&e1));
  forced.on = true;
  forced.k = Key{ family, d->dim, d->ns[0], dtype };
  forced.variant = variant;
  int rc = launch_dft(pl, dtype, dtype, dtype, f, u, out, false, s);
  hipError_t e = hipSuccess;
  if (!rc) e = hipEventRecord(e0, s);
  for (int r = 0; r < reps && !rc && e == hipSuccess; r++) rc = launch_dft(pl, dtype, dtype, dtype, f, u, out, false, s);
  if (!rc && e == hipSuccess) e = hipEventRecord(e1, s);
  if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
  forced.on = false;
  float t = 0.f;
  if (!rc && e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  if (rc) return rc;
  if (e != hipSuccess) {
    set_error(std::string("te_tune_time_dft: ") + hipGetErrorString(e));
    return TE_ERR_HIP;
  }
  *ms = (double)t / reps;
  return TE_OK;
}

}  // extern "C"
