// te_timer: the reference's Timer (Timer.h:28-204, Timer.cpp:160-781).  A tree of
// named timings, each keyed under its parent by (name, domain id, patch id), with the
// number of calls, the sum, minimum and maximum of the durations, and named int or
// double infos; serialized in the reference's JSON shape (objects with sorted keys,
// "rank" on every timing, "comm_size", the added domains as [id, patches] pairs) and
// printed in its text layout (sibling timings merged by name).
//
// A timing started on a stream is timed by two HIP events recorded on the stream and
// read when the timer is serialized or printed, so timing device work adds no
// synchronization; a timing started without a stream uses the host's steady clock, as
// the reference does.  One process: "rank" is 0 and "comm_size" 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <list>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "te_internal.h"

#define TE_TRY(x)        \
  do {                   \
    int rc_ = (x);       \
    if (rc_) return rc_; \
  } while (0)

namespace te {
namespace {

constexpr int kNone = INT_MAX;  // no domain / no patch (Timer.cpp: numeric_limits<int>::max())

// A JSON value.  Objects keep their keys sorted, as the reference's JSON library does.
struct Json {
  enum Kind { NUL, INT, DBL, STR, ARR, OBJ };
  Kind k = NUL;
  long long i = 0;
  double d = 0.0;
  std::string s;
  std::vector<Json> a;
  std::map<std::string, Json> o;

  static Json integer(long long v)
  {
    Json j;
    j.k = INT;
    j.i = v;
    return j;
  }
  static Json dbl(double v)
  {
    Json j;
    j.k = DBL;
    j.d = v;
    return j;
  }
  static Json str(const std::string& v)
  {
    Json j;
    j.k = STR;
    j.s = v;
    return j;
  }
  static Json array()
  {
    Json j;
    j.k = ARR;
    return j;
  }
  Json& operator[](const std::string& key)
  {
    if (k == NUL) k = OBJ;
    return o[key];
  }
  const Json* find(const std::string& key) const
  {
    if (k != OBJ) return nullptr;
    auto it = o.find(key);
    return it == o.end() ? nullptr : &it->second;
  }
  void push_back(Json v)
  {
    if (k == NUL) k = ARR;
    a.push_back(std::move(v));
  }
  size_t size() const { return k == ARR ? a.size() : k == OBJ ? o.size() : k == NUL ? 0 : 1; }
  double num() const { return k == INT ? (double)i : d; }
};

// shortest decimal form that reads back to the same double; integral values keep ".0"
std::string fmt_double(double v)
{
  if (!std::isfinite(v)) return "null";
  char b[40];
  for (int p = 15; p <= 17; p++) {
    snprintf(b, sizeof(b), "%.*g", p, v);
    if (std::strtod(b, nullptr) == v) break;
  }
  std::string s(b);
  if (s.find_first_of(".en") == std::string::npos) s += ".0";
  return s;
}

void dump_str(const std::string& s, std::string& out)
{
  out += '"';
  for (unsigned char c : s) {
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\b': out += "\\b"; break;
      case '\f': out += "\\f"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      default:
        if (c < 0x20) {
          char b[8];
          snprintf(b, sizeof(b), "\\u%04x", c);
          out += b;
        } else {
          out += (char)c;
        }
    }
  }
  out += '"';
}

// indent < 0: compact (no whitespace); else pretty-printed with that many spaces per level
void dump(const Json& j, std::string& out, int indent, int level)
{
  const bool pretty = indent >= 0;
  auto nl = [&](int lv) {
    if (pretty) {
      out += '\n';
      out.append((size_t)(indent * lv), ' ');
    }
  };
  switch (j.k) {
    case Json::NUL: out += "null"; break;
    case Json::INT: out += std::to_string(j.i); break;
    case Json::DBL: out += fmt_double(j.d); break;
    case Json::STR: dump_str(j.s, out); break;
    case Json::ARR:
      if (j.a.empty()) {
        out += "[]";
        break;
      }
      out += '[';
      for (size_t q = 0; q < j.a.size(); q++) {
        if (q) out += ',';
        nl(level + 1);
        dump(j.a[q], out, indent, level + 1);
      }
      nl(level);
      out += ']';
      break;
    case Json::OBJ: {
      if (j.o.empty()) {
        out += "{}";
        break;
      }
      out += '{';
      bool first = true;
      for (const auto& kv : j.o) {
        if (!first) out += ',';
        first = false;
        nl(level + 1);
        dump_str(kv.first, out);
        out += pretty ? ": " : ":";
        dump(kv.second, out, indent, level + 1);
      }
      nl(level);
      out += '}';
      break;
    }
  }
}

// ---- the domain's patches in the reference's mesh schema (json.cpp:26-247) ----

const char* const kSide[6] = { "WEST", "EAST", "SOUTH", "NORTH", "BOTTOM", "TOP" };
const char* const kEdge[12] = { "BS", "BN", "TS", "TN", "BW", "BE", "TW", "TE", "SW", "SE", "NW", "NE" };
const char* const kOrth1[2] = { "LOWER", "UPPER" };
const char* const kOrth2[4] = { "SW", "SE", "NW", "NE" };
const char* const kOrth3[8] = { "BSW", "BSE", "BNW", "BNE", "TSW", "TSE", "TNW", "TNE" };

// Orthant<m> as JSON: its name, or null (Orthant<0>, and the null orthant)
Json orthant_json(int m, int o)
{
  if (o < 0 || m == 0) return Json();
  const char* const* names = m == 1 ? kOrth1 : m == 2 ? kOrth2 : kOrth3;
  if (o >= (1 << m)) return Json();
  return Json::str(names[o]);
}

// NormalNbrInfo / CoarseNbrInfo / FineNbrInfo of slot `slot`, across a face of dimension m
Json nbr_json(const te_patch_info& p, int slot, int m)
{
  Json n;
  const int t = p.nbr_type[slot];
  const int cnt = t == 2 ? (1 << m) : 1;
  Json ids = Json::array(), ranks = Json::array();
  for (int q = 0; q < cnt; q++) {
    ids.push_back(Json::integer(p.nbr_ids[slot][q]));
    ranks.push_back(Json::integer(p.nbr_ranks[slot][q]));
  }
  n["ids"] = ids;
  n["ranks"] = ranks;
  n["type"] = Json::str(t == 0 ? "NORMAL" : t == 1 ? "COARSE" : "FINE");
  if (t == 1) n["orth_on_coarse"] = orthant_json(m, p.nbr_orth[slot]);
  return n;
}

Json patch_json(const te_patch_info& p, int dim, const int* ns)
{
  Json j;
  j["id"] = Json::integer(p.id);
  j["parent_id"] = Json::integer(p.parent_id);
  j["parent_rank"] = Json::integer(p.parent_rank);
  j["orth_on_parent"] = orthant_json(dim, p.orth_on_parent);
  j["rank"] = Json::integer(p.rank);
  Json starts = Json::array(), lengths = Json::array();
  for (int a = 0; a < dim; a++) {
    starts.push_back(Json::dbl(p.starts[a]));
    lengths.push_back(Json::dbl(p.spacings[a] * ns[a]));
  }
  j["starts"] = starts;
  j["lengths"] = lengths;
  j["refine_level"] = Json::integer(p.refine_level);
  // slots: Face<D,M>::sum_of_faces + face index (2D: corners 0-3, sides 4-7;
  // 3D: corners 0-7, edges 8-19, sides 20-25)
  const int side0 = dim == 2 ? 4 : 20;
  Json nbrs = Json::array();
  for (int s = 0; s < 2 * dim; s++) {
    if (p.nbr_type[side0 + s] < 0) continue;
    Json n = nbr_json(p, side0 + s, dim - 1);
    n["side"] = Json::str(kSide[s]);
    nbrs.push_back(n);
  }
  j["nbrs"] = nbrs;
  if (dim == 3) {
    Json en = Json::array();
    for (int e = 0; e < 12; e++) {
      if (p.nbr_type[8 + e] < 0) continue;
      Json n = nbr_json(p, 8 + e, 1);
      n["edge"] = Json::str(kEdge[e]);
      en.push_back(n);
    }
    j["edge_nbrs"] = en;
  }
  Json cn = Json::array();
  for (int c = 0; c < (1 << dim); c++) {
    if (p.nbr_type[c] < 0) continue;
    Json n = nbr_json(p, c, 0);
    n["corner"] = Json::str(dim == 2 ? kOrth2[c] : kOrth3[c]);
    cn.push_back(n);
  }
  j["corner_nbrs"] = cn;
  if (p.child_ids[0] != -1) {
    Json ci = Json::array(), cr = Json::array();
    for (int c = 0; c < (1 << dim); c++) {
      ci.push_back(Json::integer(p.child_ids[c]));
      cr.push_back(Json::integer(p.child_ranks[c]));
    }
    j["child_ids"] = ci;
    j["child_ranks"] = cr;
  }
  return j;
}

// ---- timings ----

struct Info {
  std::string name;
  bool is_int = true;
  long long isum = 0;
  int imin = INT_MAX, imax = INT_MIN;
  double dsum = 0.0, dmin = DBL_MAX, dmax = -DBL_MAX;
  int num_calls = 0;
};

struct Timing {
  Timing* parent = nullptr;
  std::string name;
  int domain_id = kNone, patch_id = kNone;
  size_t num_calls = 0;
  double sum = 0.0, max = -DBL_MAX, min = DBL_MAX;
  std::list<Timing> timings;  // children in the order first started (stable addresses)
  std::map<std::tuple<std::string, int, int>, Timing*> timing_map;
  std::vector<Info> infos;    // in the order first added
  std::map<std::string, size_t> info_map;

  Timing& child(const std::string& n, int dom, int patch)
  {
    auto key = std::make_tuple(n, dom, patch);
    auto it = timing_map.find(key);
    if (it != timing_map.end()) return *it->second;
    timings.emplace_back();
    Timing& t = timings.back();
    t.parent = this;
    t.name = n;
    t.domain_id = dom;
    t.patch_id = patch;
    timing_map.emplace(key, &t);
    return t;
  }
  void add(double seconds)
  {
    sum += seconds;
    max = std::max(max, seconds);
    min = std::min(min, seconds);
    num_calls++;
  }
};

Json info_json(const Info& f)
{
  Json j;
  j["name"] = Json::str(f.name);
  j["num_calls"] = Json::integer(f.num_calls);
  if (f.is_int) {
    j["sum"] = Json::integer(f.isum);
    j["min"] = Json::integer(f.imin);
    j["max"] = Json::integer(f.imax);
  } else {
    j["sum"] = Json::dbl(f.dsum);
    j["min"] = Json::dbl(f.dmin);
    j["max"] = Json::dbl(f.dmax);
  }
  return j;
}

// Timing's to_json (Timer.cpp:306-331) with the rank of DecorateTimingsWithRank
// (:642-660) on every timing below the root
Json timing_json(const Timing& t, bool is_root)
{
  Json j;
  if (!t.name.empty()) j["name"] = Json::str(t.name);
  if (t.domain_id != kNone) j["domain_id"] = Json::integer(t.domain_id);
  if (t.patch_id != kNone) j["patch_id"] = Json::integer(t.patch_id);
  if (t.num_calls != 0) {
    j["num_calls"] = Json::integer((long long)t.num_calls);
    j["sum"] = Json::dbl(t.sum);
    j["max"] = Json::dbl(t.max);
    j["min"] = Json::dbl(t.min);
  }
  if (!t.timings.empty()) {
    Json arr = Json::array();
    for (const Timing& c : t.timings) arr.push_back(timing_json(c, false));
    j["timings"] = arr;
  }
  if (!t.infos.empty()) {
    Json arr = Json::array();
    for (const Info& f : t.infos) arr.push_back(info_json(f));
    j["infos"] = arr;
  }
  if (!is_root) j["rank"] = Json::integer(0);
  return j;
}

// ---- the text layout of operator<< (Timer.cpp:420-640) ----

void output_line(std::ostream& os, size_t width, const std::string& name, double value)
{
  os.width((std::streamsize)(width + 1));
  os << std::right << name << std::left;
  os.width(0);
  os << ": " << value << std::endl;
}

double field(const Json& j, const char* key)
{
  const Json* v = j.find(key);
  return v ? v->num() : 0.0;
}

void merge_info(Json& a, const Json& b)
{
  a["num_calls"] = Json::integer((long long)(field(a, "num_calls") + field(b, "num_calls")));
  a["sum"] = Json::dbl(field(a, "sum") + field(b, "sum"));
  a["min"] = Json::dbl(std::min(field(a, "min"), field(b, "min")));
  a["max"] = Json::dbl(std::max(field(a, "max"), field(b, "max")));
}

void merge_timing(Json& a, const Json& b)
{
  merge_info(a, b);
  Json& ai = a["infos"];
  if (const Json* bi = b.find("infos")) {
    std::map<std::string, size_t> at;
    for (size_t q = 0; q < ai.a.size(); q++) at.emplace(ai.a[q].find("name")->s, q);
    for (const Json& f : bi->a) {
      auto ins = at.emplace(f.find("name")->s, ai.a.size());
      if (ins.second) ai.push_back(f);
      else merge_info(ai.a[ins.first->second], f);
    }
  }
  Json& at = a["timings"];
  if (const Json* bt = b.find("timings"))
    for (const Json& c : bt->a) at.push_back(c);
}

// sibling timings with the same name merged into one, in the order of first appearance
std::vector<Json> merge_timings(const Json* timings)
{
  std::vector<Json> merged;
  if (!timings) return merged;
  std::map<std::string, size_t> at;
  for (const Json& t : timings->a) {
    const Json* n = t.find("name");
    auto ins = at.emplace(n ? n->s : std::string(), merged.size());
    if (ins.second) merged.push_back(t);
    else merge_timing(merged[ins.first->second], t);
  }
  return merged;
}

void print_merged(std::ostream& os, size_t width, const std::string& parent, const Json* timings);

void print_timing(std::ostream& os, size_t width, const std::string& parent, const Json& t)
{
  const Json* n = t.find("name");
  std::string mine = parent + (n ? n->s : std::string());
  os << mine << std::endl;
  os << std::string(mine.size(), '-') << std::endl;
  const double calls = field(t, "num_calls");
  if (calls == 1.0) {
    output_line(os, width, "time (sec)", field(t, "sum"));
  } else {
    output_line(os, width, "average calls per rank", calls / 1.0);
    output_line(os, width, "average (sec)", field(t, "sum") / calls);
    output_line(os, width, "min (sec)", field(t, "min"));
    output_line(os, width, "max (sec)", field(t, "max"));
  }
  if (const Json* infos = t.find("infos")) {
    for (const Json& f : infos->a) {
      const std::string& fn = f.find("name")->s;
      const double fc = field(f, "num_calls");
      if (fc == 1.0) {
        output_line(os, width, fn, field(f, "sum"));
      } else {
        output_line(os, width, fn + " avg", field(f, "sum") / fc);
        output_line(os, width, fn + " min", field(f, "min"));
        output_line(os, width, fn + " max", field(f, "max"));
      }
    }
  }
  os << std::endl;
  print_merged(os, width, mine + " -> ", t.find("timings"));
}

void print_merged(std::ostream& os, size_t width, const std::string& parent, const Json* timings)
{
  for (const Json& t : merge_timings(timings)) print_timing(os, width, parent, t);
}

void max_info_name(const Json* timings, size_t& width)
{
  if (!timings) return;
  for (const Json& t : timings->a) {
    if (const Json* infos = t.find("infos"))
      for (const Json& f : infos->a) width = std::max(width, f.find("name")->s.size() + 4);
    max_info_name(t.find("timings"), width);
  }
}

struct Frame {
  Timing* t;
  bool on_stream;
  hipEvent_t ev;
  std::chrono::steady_clock::time_point t0;
};

struct Pending {
  Timing* t;
  hipEvent_t a, b;
};

}  // namespace
}  // namespace te

struct te_timer {
  te::Timing root;
  std::vector<te::Frame> stack;  // stack[0] is the root
  std::map<int, te::Json> domains;
  std::vector<te::Pending> pending;  // stream timings stopped and not yet read
  std::vector<hipEvent_t> pool;
};

namespace te {
namespace {

int fail(const std::string& msg)
{
  set_error(msg);
  return TE_ERR_ARG;
}

int get_event(te_timer* t, hipEvent_t* ev)
{
  if (!t->pool.empty()) {
    *ev = t->pool.back();
    t->pool.pop_back();
    return TE_OK;
  }
  hipError_t e = hipEventCreate(ev);
  if (e != hipSuccess) {
    set_error(std::string("timer event: ") + hipGetErrorString(e));
    return TE_ERR_HIP;
  }
  return TE_OK;
}

// the durations of the stream timings stopped so far, in the order they were stopped
int resolve(te_timer* t)
{
  for (size_t q = 0; q < t->pending.size(); q++) {
    Pending& p = t->pending[q];
    hipError_t e = hipEventSynchronize(p.b);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.a, p.b);
    if (e != hipSuccess) {
      t->pending.erase(t->pending.begin(), t->pending.begin() + (long)q);
      set_error(std::string("timer: reading a stream timing: ") + hipGetErrorString(e));
      return TE_ERR_HIP;
    }
    p.t->add(ms * 1e-3);
    t->pool.push_back(p.a);
    t->pool.push_back(p.b);
  }
  t->pending.clear();
  return TE_OK;
}

int start(te_timer* t, const char* name, int domain_id, int patch_id, bool on_stream, hipStream_t s)
{
  if (!t || !name) return fail("timer start: bad argument");
  if (domain_id != kNone && !t->domains.count(domain_id))
    return fail("Domain with id " + std::to_string(domain_id) + " was not added to timer");
  Timing& next = t->stack.back().t->child(name, domain_id, patch_id);
  Frame f{ &next, on_stream, nullptr, {} };
  if (on_stream) {
    TE_TRY(get_event(t, &f.ev));
    hipError_t e = hipEventRecord(f.ev, s);
    if (e != hipSuccess) {
      t->pool.push_back(f.ev);
      set_error(std::string("timer start: ") + hipGetErrorString(e));
      return TE_ERR_HIP;
    }
  } else {
    f.t0 = std::chrono::steady_clock::now();
  }
  t->stack.push_back(f);
  return TE_OK;
}

int stop(te_timer* t, const char* name, int domain_id, int patch_id, bool on_stream, hipStream_t s)
{
  if (!t || !name) return fail("timer stop: bad argument");
  Frame& f = t->stack.back();
  if (!(t->stack.size() > 1 && f.t->patch_id == patch_id && f.t->domain_id == domain_id && f.t->name == name))
    return fail("Timer was expecting to end \"" + f.t->name + "\", instead got \"" + name + "\"");
  if (f.on_stream != on_stream)
    return fail("timing \"" + f.t->name + "\" was started " + (f.on_stream ? "on a stream" : "on the host clock") +
                " and must be stopped the same way");
  if (on_stream) {
    hipEvent_t ev = nullptr;
    TE_TRY(get_event(t, &ev));
    hipError_t e = hipEventRecord(ev, s);
    if (e != hipSuccess) {
      t->pool.push_back(ev);
      set_error(std::string("timer stop: ") + hipGetErrorString(e));
      return TE_ERR_HIP;
    }
    t->pending.push_back(Pending{ f.t, f.ev, ev });
    // keep the number of events in flight bounded
    if (t->pending.size() >= 4096) TE_TRY(resolve(t));
  } else {
    f.t->add(std::chrono::duration<double>(std::chrono::steady_clock::now() - f.t0).count());
  }
  t->stack.pop_back();
  return TE_OK;
}

int add_info(te_timer* t, const char* name, bool is_int, int iv, double dv)
{
  if (!t || !name) return fail("timer info: bad argument");
  if (t->stack.size() == 1) return fail("No timing to add information to");
  Timing& cur = *t->stack.back().t;
  auto ins = cur.info_map.emplace(name, cur.infos.size());
  if (ins.second) {
    cur.infos.emplace_back();
    cur.infos.back().name = name;
    cur.infos.back().is_int = is_int;
  }
  Info& f = cur.infos[ins.first->second];
  if (f.is_int != is_int)
    return fail(std::string("Adding ") + (is_int ? "int" : "double") + " info to existing " +
                (is_int ? "double" : "int") + " info of timing " + cur.name + " for info " + name);
  if (is_int) {
    f.isum += iv;
    f.imin = std::min(f.imin, iv);
    f.imax = std::max(f.imax, iv);
  } else {
    f.dsum += dv;
    f.dmin = std::min(f.dmin, dv);
    f.dmax = std::max(f.dmax, dv);
  }
  f.num_calls++;
  return TE_OK;
}

// to_json(json&, const Timer&) (Timer.cpp:700-745) for one rank
int timer_json(te_timer* t, Json* out)
{
  TE_TRY(resolve(t));
  Json j = timing_json(t->root, true);
  if (!t->domains.empty()) {
    Json arr = Json::array();
    for (const auto& kv : t->domains) {
      Json pair = Json::array();
      pair.push_back(Json::integer(kv.first));
      pair.push_back(kv.second);
      arr.push_back(pair);
    }
    j["domains"] = arr;
  }
  if (j.k != Json::NUL) j["comm_size"] = Json::integer(1);
  *out = std::move(j);
  return TE_OK;
}

int copy_out(const std::string& s, char* buf, size_t cap, size_t* len)
{
  if (len) *len = s.size();
  if (!buf) return TE_OK;
  if (cap <= s.size()) return fail("timer: buffer of " + std::to_string(cap) + " bytes is too small (" +
                                   std::to_string(s.size() + 1) + " needed)");
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return TE_OK;
}

}  // namespace

int timer_start_stream(te_timer* t, const char* name, int domain_id, hipStream_t s)
{
  return t ? start(t, name, domain_id, kNone, true, s) : TE_OK;
}

int timer_stop_stream(te_timer* t, const char* name, int domain_id, hipStream_t s)
{
  return t ? stop(t, name, domain_id, kNone, true, s) : TE_OK;
}

bool timer_has_domain(const te_timer* t, int domain_id) { return t && t->domains.count(domain_id) > 0; }

}  // namespace te

using namespace te;

extern "C" {

int te_timer_create(te_timer** out)
{
  if (!out) return fail("te_timer_create: bad argument");
  te_timer* t = new te_timer();
  t->stack.push_back(Frame{ &t->root, false, nullptr, {} });
  *out = t;
  return TE_OK;
}

int te_timer_destroy(te_timer* t)
{
  if (!t) return TE_OK;
  for (const Pending& p : t->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (const Frame& f : t->stack)
    if (f.on_stream && f.ev) (void)hipEventDestroy(f.ev);
  for (hipEvent_t e : t->pool) (void)hipEventDestroy(e);
  delete t;
  return TE_OK;
}

int te_timer_add_domain(te_timer* t, int domain_id, int dim, const int* ns, int num_patches,
                        const te_patch_info* patches)
{
  if (!t || (dim != 2 && dim != 3) || !ns || num_patches < 0 || (num_patches > 0 && !patches))
    return fail("te_timer_add_domain: bad argument");
  if (t->domains.count(domain_id))
    return fail("Domain with id " + std::to_string(domain_id) + " was already added to timer");
  Json j;  // null for a domain without patches, as the reference's to_json(Domain) leaves it
  for (int q = 0; q < num_patches; q++) j.push_back(patch_json(patches[q], dim, ns));
  t->domains.emplace(domain_id, std::move(j));
  return TE_OK;
}

int te_timer_start(te_timer* t, const char* name, int domain_id, int patch_id)
{
  return start(t, name, domain_id, patch_id, false, nullptr);
}

int te_timer_stop(te_timer* t, const char* name, int domain_id, int patch_id)
{
  return stop(t, name, domain_id, patch_id, false, nullptr);
}

int te_timer_start_on(te_timer* t, const char* name, int domain_id, int patch_id, te_stream stream)
{
  return start(t, name, domain_id, patch_id, true, (hipStream_t)stream);
}

int te_timer_stop_on(te_timer* t, const char* name, int domain_id, int patch_id, te_stream stream)
{
  return stop(t, name, domain_id, patch_id, true, (hipStream_t)stream);
}

int te_timer_add_int_info(te_timer* t, const char* name, int info) { return add_info(t, name, true, info, 0.0); }

int te_timer_add_double_info(te_timer* t, const char* name, double info)
{
  return add_info(t, name, false, 0, info);
}

int te_timer_to_json(te_timer* t, int indent, char* buf, size_t cap, size_t* len)
{
  if (!t) return fail("te_timer_to_json: bad argument");
  Json j;
  TE_TRY(timer_json(t, &j));
  std::string s;
  dump(j, s, indent, 0);
  return copy_out(s, buf, cap, len);
}

int te_timer_print(te_timer* t, char* buf, size_t cap, size_t* len)
{
  if (!t) return fail("te_timer_print: bad argument");
  if (t->stack.size() > 1)
    return fail("Cannot output Timer results with unfinished timings, check that all timings have been stopped. "
                "Currently waiting on timing \"" + t->stack.back().t->name + "\"");
  Json j;
  TE_TRY(timer_json(t, &j));
  std::ostringstream os;
  os << std::endl << "TIMING RESULTS" << std::endl << "==============" << std::endl << std::endl;
  const Json* timings = j.find("timings");
  if (!timings) {
    os << "No timings to report." << std::endl << std::endl;
  } else {
    size_t width = std::strlen("average calls per rank");
    max_info_name(timings, width);
    print_merged(os, width, "", timings);
  }
  return copy_out(os.str(), buf, cap, len);
}

int te_timer_save(te_timer* t, const char* filename)
{
  if (!t || !filename) return fail("te_timer_save: bad argument");
  Json j;
  TE_TRY(timer_json(t, &j));
  std::string s;
  dump(j, s, 4, 0);
  std::ofstream out(filename, std::ofstream::out | std::ofstream::trunc);
  if (out.fail()) return fail(std::string("Failed to open file ") + filename);
  out << s;
  out.close();
  return TE_OK;
}

}  // extern "C"
