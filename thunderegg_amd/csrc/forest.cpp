// Quadtree / octree forest on the unit square / cube and the Domain tables of
// every multigrid level, replacing P4estDomainGenerator / P8estDomainGenerator
// (P4estDomainGenerator.cpp:205-738, P8estDomainGenerator.cpp:812-852).
//
// Semantics kept from the reference generators:
//  - leaves in Morton order (child order = Orthant bits x, y, z); the finest
//    domain's ids are the leaf positions (SetIds, :180-198);
//  - coarser domains are made one level at a time by merging complete families
//    at the current finest level (coarsenTree, :271-322); a merged patch takes
//    the id of its first (SW) child, children record parent_id / orth_on_parent,
//    unmerged patches are their own parent with child_ids = {id, -1, ...};
//  - neighbours: faces (Normal / Coarse with orth_on_coarse / Fine with the
//    2^(D-1) ids in face-orthant order, :384-474), corners only where the corner
//    point is a corner of every leaf around it (p4est visits no hanging
//    corners), 3D edges where the edge is a full edge or the half of a full edge
//    of the leaves around it (:812-852).
// One rank: rank = parent_rank = 0.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "te_internal.h"

using namespace te;

namespace {

constexpr int kMaxLevel = 19;

struct Cell {
  int l;
  int c[3];
};

inline uint64_t key_of(int l, const int* c)
{
  return ((uint64_t)l << 57) | ((uint64_t)(uint32_t)c[0] << 38) | ((uint64_t)(uint32_t)c[1] << 19) |
         (uint64_t)(uint32_t)c[2];
}

// Morton key at the finest resolution (interleaved bits, x lowest)
inline uint64_t morton(int dim, const Cell& q)
{
  uint64_t m = 0;
  int sh = kMaxLevel - q.l;
  for (int b = kMaxLevel - 1; b >= 0; b--) {
    for (int a = dim - 1; a >= 0; a--) {
      m = (m << 1) | (((uint64_t)q.c[a] << sh) >> b & 1);
    }
  }
  return m;
}

struct Level {
  std::vector<Cell> cells;                   // Morton order
  std::vector<int> ids;
  std::unordered_map<uint64_t, int> index;   // key -> position
  std::vector<te_patch_info> patches;
};

}  // namespace

struct te_forest {
  int dim = 2;
  std::vector<Cell> leaves;  // Morton order
  std::vector<Level> domains;  // finest first
};

namespace {

void sort_morton(int dim, std::vector<Cell>& v)
{
  std::vector<std::pair<uint64_t, Cell>> tmp;
  tmp.reserve(v.size());
  for (auto& q : v) tmp.emplace_back(morton(dim, q), q);
  std::sort(tmp.begin(), tmp.end(), [](const auto& a, const auto& b) {
    return a.first < b.first || (a.first == b.first && a.second.l < b.second.l);
  });
  for (size_t i = 0; i < v.size(); i++) v[i] = tmp[i].second;
}

struct Finder {
  int dim;
  const std::unordered_map<uint64_t, int>* idx;
  bool exists(int l, const int* c) const
  {
    if (l < 0) return false;
    for (int a = 0; a < dim; a++)
      if (c[a] < 0 || c[a] >= (1 << l)) return false;
    int cc[3] = { c[0], c[1], dim == 3 ? c[2] : 0 };
    return idx->count(key_of(l, cc)) != 0;
  }
  int get(int l, const int* c) const
  {
    int cc[3] = { c[0], c[1], dim == 3 ? c[2] : 0 };
    auto it = idx->find(key_of(l, cc));
    return it == idx->end() ? -1 : it->second;
  }
  // leaf (index) containing the cell (l, c), searching coarser levels; -1 if none
  int containing(int l, const int* c, int* lev) const
  {
    for (int a = 0; a < dim; a++)
      if (c[a] < 0 || c[a] >= (1 << l)) return -1;
    for (int ll = l; ll >= 0; ll--) {
      int cc[3] = { 0, 0, 0 };
      for (int a = 0; a < dim; a++) cc[a] = c[a] >> (l - ll);
      auto it = idx->find(key_of(ll, cc));
      if (it != idx->end()) {
        if (lev) *lev = ll;
        return it->second;
      }
    }
    return -1;
  }
};

void set_nbr(te_patch_info& p, int slot, int type, int orth, const std::vector<int>& ids)
{
  p.nbr_type[slot] = type;
  p.nbr_orth[slot] = orth;
  for (size_t k = 0; k < ids.size(); k++) {
    p.nbr_ids[slot][k] = ids[k];
    p.nbr_ranks[slot][k] = 0;
  }
}

// Status of a region (cell at level l) around a line / point: 1 full leaf, 2 split into
// leaves one level finer (the caller checks which), 0 otherwise.
int link_level(int dim, Level& L, std::string* err)
{
  Finder F{ dim, &L.index };
  const int nf = dim == 2 ? 8 : 26;
  const int sb = dim == 2 ? 4 : 20;
  for (size_t i = 0; i < L.cells.size(); i++) {
    te_patch_info& P = L.patches[i];
    for (int s = 0; s < nf; s++) {
      P.nbr_type[s] = -1;
      P.nbr_orth[s] = -1;
      for (int k = 0; k < 4; k++) {
        P.nbr_ids[s][k] = -1;
        P.nbr_ranks[s][k] = -1;
      }
    }
    const Cell& q = L.cells[i];
    const int l = q.l;
    // ---- faces
    for (int s = 0; s < 2 * dim; s++) {
      int a = s >> 1, up = s & 1;
      int c2[3] = { q.c[0], q.c[1], q.c[2] };
      c2[a] += up ? 1 : -1;
      if (c2[a] < 0 || c2[a] >= (1 << l)) continue;
      int tang[2], nt = 0;
      for (int b = 0; b < dim; b++)
        if (b != a) tang[nt++] = b;
      if (F.exists(l, c2)) {
        set_nbr(P, sb + s, 0, -1, { L.ids[F.get(l, c2)] });
      } else if (l > 0) {
        int cp[3] = { c2[0] >> 1, c2[1] >> 1, c2[2] >> 1 };
        if (F.exists(l - 1, cp)) {
          int orth = 0;
          for (int k = 0; k < nt; k++) orth |= (q.c[tang[k]] & 1) << k;
          set_nbr(P, sb + s, 1, orth, { L.ids[F.get(l - 1, cp)] });
          continue;
        }
      }
      if (P.nbr_type[sb + s] >= 0) continue;
      std::vector<int> ids;
      for (int o = 0; o < (1 << nt); o++) {
        int cf[3] = { 2 * c2[0], 2 * c2[1], 2 * c2[2] };
        cf[a] += up ? 0 : 1;
        for (int k = 0; k < nt; k++) cf[tang[k]] += (o >> k) & 1;
        int j = F.get(l + 1, cf);
        if (j < 0) {
          *err = "forest is not 2:1 balanced across a face";
          return TE_ERR_ARG;
        }
        ids.push_back(L.ids[j]);
      }
      set_nbr(P, sb + s, 2, -1, ids);
    }
    // ---- corners: only where the corner point is a corner of all 2^D leaves around it
    for (int cr = 0; cr < (1 << dim); cr++) {
      int X[3];  // corner point in units of level l+1
      for (int a = 0; a < dim; a++) X[a] = 2 * (q.c[a] + ((cr >> a) & 1));
      bool ok = true;
      int diag = -1, diag_l = -1;
      for (int o = 0; o < (1 << dim) && ok; o++) {
        int cell[3] = { 0, 0, 0 };
        for (int a = 0; a < dim; a++) cell[a] = X[a] + (((o >> a) & 1) ? 0 : -1);
        int lv = -1;
        int j = F.containing(l + 1, cell, &lv);
        if (j < 0) {
          ok = false;
          break;
        }
        int sh = (l + 1) - lv;
        for (int a = 0; a < dim; a++)
          if (X[a] % (1 << sh)) ok = false;
        if (o == cr) {
          diag = j;
          diag_l = lv;
        }
      }
      if (!ok || diag < 0) continue;
      if (diag_l > l) set_nbr(P, cr, 2, -1, { L.ids[diag] });
      else if (diag_l < l) set_nbr(P, cr, 1, 0, { L.ids[diag] });
      else set_nbr(P, cr, 0, -1, { L.ids[diag] });
    }
    // ---- edges (3D)
    if (dim == 3) {
      for (int e = 0; e < 12; e++) {
        int fa = e >> 2;
        int fx[2], nf2 = 0;
        for (int b = 0; b < 3; b++)
          if (b != fa) fx[nf2++] = b;
        int bits[2] = { e & 1, (e >> 1) & 1 };
        // region status around the edge line at level lv, for the level-lv cell base
        // (the cell on P's side of the line); d1/d2 select across (1) or same side (0)
        auto region = [&](int lv, const int* base, int d1, int d2, int* full_idx, int* half_idx) -> int {
          int rc[3] = { base[0], base[1], base[2] };
          if (d1) rc[fx[0]] += bits[0] ? 1 : -1;
          if (d2) rc[fx[1]] += bits[1] ? 1 : -1;
          for (int a = 0; a < 3; a++)
            if (rc[a] < 0 || rc[a] >= (1 << lv)) return -1;  // outside the domain
          if (F.exists(lv, rc)) {
            *full_idx = F.get(lv, rc);
            return 1;
          }
          // the two level lv+1 cells along the line inside the region
          for (int h = 0; h < 2; h++) {
            int cf[3] = { 2 * rc[0], 2 * rc[1], 2 * rc[2] };
            // adjacent to the line: on the side facing P's cell
            cf[fx[0]] += d1 ? (bits[0] ? 0 : 1) : (bits[0] ? 1 : 0);
            cf[fx[1]] += d2 ? (bits[1] ? 0 : 1) : (bits[1] ? 1 : 0);
            cf[fa] += h;
            int j = F.get(lv + 1, cf);
            if (j < 0) return 0;
            half_idx[h] = j;
          }
          return 2;
        };
        auto visited = [&](int lv, const int* base, int st[4], int full[4], int half[4][2]) -> bool {
          int nfull = 0;
          for (int r = 0; r < 4; r++) {
            st[r] = region(lv, base, r & 1, r >> 1, &full[r], half[r]);
            if (st[r] <= 0) return false;
            if (st[r] == 1) nfull++;
          }
          return nfull > 0;
        };
        int st[4], full[4], half[4][2];
        int base[3] = { q.c[0], q.c[1], q.c[2] };
        int slot = 8 + e;
        if (visited(l, base, st, full, half) && st[0] == 1) {
          if (st[3] == 1) set_nbr(P, slot, 0, -1, { L.ids[full[3]] });
          else set_nbr(P, slot, 2, -1, { L.ids[half[3][0]], L.ids[half[3][1]] });
          continue;
        }
        // P as one half of a full edge one level coarser
        if (l == 0) continue;
        if ((q.c[fx[0]] & 1) != bits[0] || (q.c[fx[1]] & 1) != bits[1]) continue;
        int pb[3] = { q.c[0] >> 1, q.c[1] >> 1, q.c[2] >> 1 };
        if (!visited(l - 1, pb, st, full, half) || st[0] != 2) continue;
        int pos = q.c[fa] & 1;
        if (st[3] == 1) set_nbr(P, slot, 1, pos, { L.ids[full[3]] });
        else set_nbr(P, slot, 0, -1, { L.ids[half[3][pos]] });  // both diagonal sides hanging
      }
    }
  }
  return TE_OK;
}

}  // namespace

extern "C" {

int te_forest_create(int dim, int base_level, te_forest** out)
{
  if ((dim != 2 && dim != 3) || base_level < 0 || base_level > 10 || !out) {
    set_error("te_forest_create: bad argument");
    return TE_ERR_ARG;
  }
  auto f = new te_forest();
  f->dim = dim;
  int n = 1 << base_level;
  int nz = dim == 3 ? n : 1;
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < n; y++)
      for (int x = 0; x < n; x++) f->leaves.push_back(Cell{ base_level, { x, y, z } });
  sort_morton(dim, f->leaves);
  *out = f;
  return TE_OK;
}

int te_forest_destroy(te_forest* f)
{
  delete f;
  return TE_OK;
}

int te_forest_num_leaves(const te_forest* f) { return f ? (int)f->leaves.size() : -1; }

int te_forest_get_leaves(const te_forest* f, int32_t* levels, int32_t* coords)
{
  if (!f) return TE_ERR_ARG;
  for (size_t i = 0; i < f->leaves.size(); i++) {
    if (levels) levels[i] = f->leaves[i].l;
    if (coords)
      for (int a = 0; a < f->dim; a++) coords[i * f->dim + a] = f->leaves[i].c[a];
  }
  return TE_OK;
}

int te_forest_refine(te_forest* f, const uint8_t* flags)
{
  if (!f || !flags) return TE_ERR_ARG;
  std::vector<Cell> nl;
  nl.reserve(f->leaves.size() * 2);
  for (size_t i = 0; i < f->leaves.size(); i++) {
    const Cell& q = f->leaves[i];
    if (flags[i] && q.l < kMaxLevel - 1) {
      for (int o = 0; o < (1 << f->dim); o++) {
        Cell ch{ q.l + 1, { 2 * q.c[0] + (o & 1), 2 * q.c[1] + ((o >> 1) & 1), f->dim == 3 ? 2 * q.c[2] + ((o >> 2) & 1) : 0 } };
        nl.push_back(ch);
      }
    } else {
      nl.push_back(q);
    }
  }
  sort_morton(f->dim, nl);
  f->leaves.swap(nl);
  f->domains.clear();
  return TE_OK;
}

// 2:1 balance over faces, edges and corners: refine any leaf that has a
// neighbour (in any of the 3^D - 1 directions) more than one level finer.
int te_forest_balance(te_forest* f)
{
  if (!f) return TE_ERR_ARG;
  int dim = f->dim;
  for (int iter = 0; iter < 64; iter++) {
    std::unordered_map<uint64_t, int> idx;
    for (size_t i = 0; i < f->leaves.size(); i++) idx[key_of(f->leaves[i].l, f->leaves[i].c)] = (int)i;
    Finder F{ dim, &idx };
    std::vector<uint8_t> flag(f->leaves.size(), 0);
    bool any = false;
    int nd = 1;
    for (int a = 0; a < dim; a++) nd *= 3;
    for (size_t i = 0; i < f->leaves.size(); i++) {
      const Cell& q = f->leaves[i];
      for (int dd = 0; dd < nd; dd++) {
        int off[3] = { dd % 3 - 1, (dd / 3) % 3 - 1, dim == 3 ? dd / 9 - 1 : 0 };
        if (off[0] == 0 && off[1] == 0 && off[2] == 0) continue;
        int c2[3] = { q.c[0] + off[0], q.c[1] + off[1], q.c[2] + off[2] };
        int lv = -1;
        int j = F.containing(q.l, c2, &lv);
        if (j >= 0 && lv < q.l - 1 && !flag[j]) {
          flag[j] = 1;
          any = true;
        }
      }
    }
    if (!any) return TE_OK;
    int rc = te_forest_refine(f, flag.data());
    if (rc) return rc;
  }
  set_error("te_forest_balance: did not converge");
  return TE_ERR_ARG;
}

int te_forest_build_domains(te_forest* f, const int* ns, int* num_levels)
{
  if (!f || !ns || !num_levels) return TE_ERR_ARG;
  int dim = f->dim;
  int lmax = 0;
  for (auto& q : f->leaves) lmax = std::max(lmax, q.l);
  f->domains.assign(lmax + 1, Level());
  // finest domain: the leaves, ids = Morton position
  {
    Level& L = f->domains[0];
    L.cells = f->leaves;
    L.ids.resize(L.cells.size());
    for (size_t i = 0; i < L.cells.size(); i++) L.ids[i] = (int)i;
  }
  // coarser domains: merge the complete families at the current finest level
  for (int k = 1; k <= lmax; k++) {
    int cur = lmax - k;  // level of the new domain's finest patches
    Level& Fn = f->domains[k - 1];
    Level& C = f->domains[k];
    for (size_t i = 0; i < Fn.cells.size(); i++) {
      const Cell& q = Fn.cells[i];
      if (q.l > cur) {
        if (((q.c[0] | q.c[1] | q.c[2]) & 1) == 0) {  // first child of its family
          C.cells.push_back(Cell{ q.l - 1, { q.c[0] >> 1, q.c[1] >> 1, q.c[2] >> 1 } });
          C.ids.push_back(Fn.ids[i]);
        }
      } else {
        C.cells.push_back(q);
        C.ids.push_back(Fn.ids[i]);
      }
    }
  }
  for (int k = 0; k <= lmax; k++) {
    Level& L = f->domains[k];
    L.index.clear();
    for (size_t i = 0; i < L.cells.size(); i++) L.index[key_of(L.cells[i].l, L.cells[i].c)] = (int)i;
    L.patches.assign(L.cells.size(), te_patch_info());
    for (size_t i = 0; i < L.cells.size(); i++) {
      te_patch_info& P = L.patches[i];
      std::memset(&P, 0, sizeof(P));
      const Cell& q = L.cells[i];
      P.id = L.ids[i];
      P.rank = 0;
      P.parent_id = -1;
      P.parent_rank = -1;
      P.orth_on_parent = -1;
      P.refine_level = q.l;
      for (int o = 0; o < 8; o++) {
        P.child_ids[o] = -1;
        P.child_ranks[o] = -1;
      }
      double h = 1.0 / (double)(1 << q.l);
      for (int a = 0; a < 3; a++) {
        P.starts[a] = a < dim ? q.c[a] * h : 0.0;
        P.spacings[a] = a < dim ? h / ns[a] : 1.0;
      }
    }
    std::string err;
    int rc = link_level(dim, L, &err);
    if (rc) {
      set_error(err);
      return rc;
    }
  }
  // parents (coarsenTree callback, :271-322) and children (replace callback)
  for (int k = 0; k < lmax; k++) {
    Level& Fn = f->domains[k];
    Level& C = f->domains[k + 1];
    Finder FC{ dim, &C.index };
    for (size_t i = 0; i < Fn.cells.size(); i++) {
      const Cell& q = Fn.cells[i];
      te_patch_info& P = Fn.patches[i];
      P.parent_rank = 0;
      int j = FC.get(q.l, q.c);
      if (j >= 0) {  // not merged
        P.parent_id = P.id;
        P.orth_on_parent = -1;
        C.patches[j].child_ids[0] = P.id;
        C.patches[j].child_ranks[0] = 0;
      } else {
        int cp[3] = { q.c[0] >> 1, q.c[1] >> 1, q.c[2] >> 1 };
        int pj = FC.get(q.l - 1, cp);
        if (pj < 0) {
          set_error("te_forest_build_domains: missing parent");
          return TE_ERR_ARG;
        }
        int orth = (q.c[0] & 1) | ((q.c[1] & 1) << 1) | (dim == 3 ? (q.c[2] & 1) << 2 : 0);
        P.parent_id = C.ids[pj];
        P.orth_on_parent = orth;
        C.patches[pj].child_ids[orth] = P.id;
        C.patches[pj].child_ranks[orth] = 0;
      }
    }
  }
  *num_levels = lmax + 1;
  return TE_OK;
}

int te_forest_domain_size(const te_forest* f, int level, int* num_patches)
{
  if (!f || level < 0 || level >= (int)f->domains.size() || !num_patches) return TE_ERR_ARG;
  *num_patches = (int)f->domains[level].patches.size();
  return TE_OK;
}

int te_forest_domain_patches(const te_forest* f, int level, te_patch_info* out)
{
  if (!f || level < 0 || level >= (int)f->domains.size() || !out) return TE_ERR_ARG;
  const auto& v = f->domains[level].patches;
  std::memcpy(out, v.data(), sizeof(te_patch_info) * v.size());
  return TE_OK;
}

}  // extern "C"
