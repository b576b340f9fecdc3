// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY: C entry points over
// oracle.hpp for the Python checkers (ctypes).  See oracle.hpp's header.
#include <chrono>
#include <cstdio>
#include <numeric>

#include "oracle.hpp"

using namespace oracle;

extern "C" {

// A body/shape as the tests describe it: type (0 sphere, 1 box, 2 capsule,
// 3 hull, 4 plane), params (sphere r | box hx,hy,hz | capsule r,hh |
// plane nx,ny,nz,c), hull vertex pointer (xyz triplets) + count.
typedef struct {
  int32_t type;
  float p[4];
  const float* hull;
  int32_t nhull;
} or_shape;

typedef struct {
  int64_t n;
  const float* pos;      // 3n
  const float* quat;     // 4n (x,y,z,w)
  const uint8_t* type;   // n
  const float* params;   // 4n
  const uint32_t* hull_first;  // n (hull bodies only)
  const uint32_t* hull_count;  // n
  const float* hull_verts;     // 3 * total
  int64_t n_hull_verts;
} or_scene;

}  // extern "C"

namespace {

template <class T>
struct HullStore {
  std::vector<V3<T>> pts;
};

template <class T>
Shape<T> make_shape(int type, const float* p, const float* hull, int nhull, std::vector<V3<T>>& store) {
  switch (type) {
    case SPHERE: return Shape<T>::sphere((T)p[0]);
    case BOX: return Shape<T>::box(mk((T)p[0], (T)p[1], (T)p[2]));
    case CAPSULE: return Shape<T>::capsule((T)p[0], (T)p[1]);
    case PLANE: return Shape<T>::plane(mk((T)p[0], (T)p[1], (T)p[2]), (T)p[3]);
    case HULL: {
      store.resize(nhull);
      for (int i = 0; i < nhull; ++i) store[i] = mk((T)hull[3 * i], (T)hull[3 * i + 1], (T)hull[3 * i + 2]);
      return Shape<T>::hull(store.data(), nhull);
    }
  }
  return Shape<T>::sphere(T(0));
}

template <class T>
Shape<T> from_desc(const or_shape* s, std::vector<V3<T>>& store) {
  return make_shape<T>(s->type, s->p, s->hull, s->nhull, store);
}

template <class T> V3<T> v3p(const float* p) { return mk((T)p[0], (T)p[1], (T)p[2]); }
template <class T> Quat<T> q4p(const float* q) { return Quat<T>{(T)q[0], (T)q[1], (T)q[2], (T)q[3]}; }
template <class T> V3<T> v3p(const double* p) { return mk((T)p[0], (T)p[1], (T)p[2]); }
template <class T> Quat<T> q4p(const double* q) { return Quat<T>{(T)q[0], (T)q[1], (T)q[2], (T)q[3]}; }
template <class T> V3<T> v3d(const double* p) { return mk((T)p[0], (T)p[1], (T)p[2]); }
template <class T> Quat<T> q4d(const double* q) { return Quat<T>{(T)q[0], (T)q[1], (T)q[2], (T)q[3]}; }

// All body shapes of a scene, with hull vertices converted once.
template <class T>
struct SceneShapes {
  std::vector<Shape<T>> shapes;
  std::vector<V3<T>> hull_pts;
  std::vector<V3<T>> pos;
  std::vector<Quat<T>> quat;
  void build(const or_scene* sc) {
    hull_pts.resize(sc->n_hull_verts);
    for (int64_t i = 0; i < sc->n_hull_verts; ++i)
      hull_pts[i] = mk((T)sc->hull_verts[3 * i], (T)sc->hull_verts[3 * i + 1], (T)sc->hull_verts[3 * i + 2]);
    shapes.resize(sc->n);
    pos.resize(sc->n);
    quat.resize(sc->n);
    for (int64_t i = 0; i < sc->n; ++i) {
      const float* p = sc->params + 4 * i;
      int t = sc->type[i];
      if (t == HULL) {
        shapes[i] = Shape<T>::hull(hull_pts.data() + sc->hull_first[i], (int)sc->hull_count[i]);
      } else {
        std::vector<V3<T>> dummy;
        shapes[i] = make_shape<T>(t, p, nullptr, 0, dummy);
      }
      pos[i] = v3p<T>(sc->pos + 3 * i);
      quat[i] = q4p<T>(sc->quat + 4 * i);
    }
  }
};

template <class T>
void write_manifold(const Manifold<T>& m, T* out) {
  for (int k = 0; k < m.n && k < 4; ++k) {
    T* o = out + 7 * k;
    o[0] = m.pts[k].position.x; o[1] = m.pts[k].position.y; o[2] = m.pts[k].position.z;
    o[3] = m.pts[k].normal.x; o[4] = m.pts[k].normal.y; o[5] = m.pts[k].normal.z;
    o[6] = m.pts[k].penetration;
  }
}

// Sort-and-sweep over x with exact closed-interval filtering: the same pair
// set as brute force (util.rs:28-38 over all i<j), ordered lexicographically.
template <class T>
int64_t sweep_pairs(int64_t n, const Aabb<T>* boxes, std::vector<uint64_t>& keys) {
  std::vector<int64_t> order;
  order.reserve(n);
  for (int64_t i = 0; i < n; ++i)
    if (!std::isnan(boxes[i].min.x)) order.push_back(i);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return boxes[a].min.x < boxes[b].min.x; });
  keys.clear();
  for (size_t k = 0; k < order.size(); ++k) {
    const Aabb<T>& a = boxes[order[k]];
    for (size_t l = k + 1; l < order.size(); ++l) {
      const Aabb<T>& b = boxes[order[l]];
      if (b.min.x > a.max.x) break;
      if (a.overlaps(b)) {
        uint64_t i = (uint64_t)order[k], j = (uint64_t)order[l];
        if (i > j) std::swap(i, j);
        keys.push_back((i << 32) | j);
      }
    }
  }
  std::sort(keys.begin(), keys.end());
  return (int64_t)keys.size();
}

// DbvtBroadPhase on a fresh scene (broadphase/dbvt.rs): create_proxy for
// slots 0..n-1 in order (:679-697) inserts each AABB into the dynamic tree
// (insert_leaf, :159-224) and runs collide_leaf (:629-645), which walks the
// tree for leaves overlapping the new one (walk_overlapping, :421-443) and
// records (other, slot).  On a fresh scene each pair is found once, when its
// later proxy arrives.  The reference keeps them in a sorted Vec by binary
// search and insertion (:593-598); here they are collected and sorted once:
// the same set in the same order, without the O(P) shifting per insertion.
// compute_pairs' tick (:647-677, one rotation path) changes no leaf volume
// on a fresh scene and its exact filter (:768-806) passes every pair, so
// both are left out.  Glam's Vec3::min / max (merged, util.rs:41-46) keep
// the non-NaN operand, as std::fmin / fmax do.
struct DbvtOracle {
  static constexpr uint32_t NONE = 0xffffffffu;
  struct Node {
    Aabb<float> vol;
    uint32_t parent, c0, c1;  // c0 == NONE: leaf, c1 = payload
  };
  std::vector<Node> nodes;
  uint32_t root = NONE;

  static Aabb<float> merged(const Aabb<float>& a, const Aabb<float>& b) {
    return Aabb<float>{mk(std::fmin(a.min.x, b.min.x), std::fmin(a.min.y, b.min.y), std::fmin(a.min.z, b.min.z)),
                       mk(std::fmax(a.max.x, b.max.x), std::fmax(a.max.y, b.max.y), std::fmax(a.max.z, b.max.z))};
  }
  // util.rs:61-68
  static bool contains(const Aabb<float>& a, const Aabb<float>& b) {
    return a.min.x <= b.min.x && a.min.y <= b.min.y && a.min.z <= b.min.z && b.max.x <= a.max.x &&
           b.max.y <= a.max.y && b.max.z <= a.max.z;
  }
  // dbvt.rs:77-80
  static float proximity(const Aabb<float>& a, const Aabb<float>& b) {
    V3<float> d = (a.min + a.max) - (b.min + b.max);
    return (std::fabs(d.x) + std::fabs(d.y)) + std::fabs(d.z);
  }
  bool leaf(uint32_t n) const { return nodes[n].c0 == NONE; }

  uint32_t insert(const Aabb<float>& v, uint32_t payload) {
    uint32_t lf = (uint32_t)nodes.size();
    nodes.push_back(Node{v, NONE, NONE, payload});
    if (root == NONE) {
      root = lf;
      return lf;
    }
    uint32_t node = root;
    while (!leaf(node)) {
      uint32_t c0 = nodes[node].c0, c1 = nodes[node].c1;
      node = proximity(v, nodes[c0].vol) < proximity(v, nodes[c1].vol) ? c0 : c1;
    }
    uint32_t prev = nodes[node].parent;
    uint32_t in = (uint32_t)nodes.size();
    nodes.push_back(Node{merged(nodes[node].vol, v), prev, node, lf});
    nodes[node].parent = in;
    nodes[lf].parent = in;
    if (prev != NONE) {
      if (nodes[prev].c1 == node) nodes[prev].c1 = in;
      else nodes[prev].c0 = in;
    } else {
      root = in;
    }
    uint32_t child = in, p = prev;
    while (p != NONE) {
      if (contains(nodes[p].vol, nodes[child].vol)) break;
      nodes[p].vol = merged(nodes[nodes[p].c0].vol, nodes[nodes[p].c1].vol);
      child = p;
      p = nodes[p].parent;
    }
    return lf;
  }

  template <class F>
  void walk(const Aabb<float>& v, F f) const {
    if (root == NONE) return;
    std::vector<uint32_t> stack{root};
    while (!stack.empty()) {
      uint32_t n = stack.back();
      stack.pop_back();
      if (!nodes[n].vol.overlaps(v)) continue;
      if (leaf(n)) {
        f(nodes[n].c1);
      } else {
        stack.push_back(nodes[n].c0);  // the right child is visited first
        stack.push_back(nodes[n].c1);
      }
    }
  }
};

template <class T>
int64_t dbvt_pairs(int64_t n, const Aabb<T>* boxes, std::vector<uint64_t>& keys) {
  DbvtOracle t;
  t.nodes.reserve(2 * (size_t)n);
  keys.clear();
  for (int64_t s = 0; s < n; ++s) {
    Aabb<float> v = boxes[s];
    t.insert(v, (uint32_t)s);
    t.walk(v, [&](uint32_t other) {
      if (other != (uint32_t)s) keys.push_back(((uint64_t)other << 32) | (uint64_t)s);
    });
  }
  std::sort(keys.begin(), keys.end());
  return (int64_t)keys.size();
}

}  // namespace

extern "C" {

float oracle_threshold_f32(const or_shape* s) {
  std::vector<V3<float>> st;
  return from_desc<float>(s, st).contact_breaking_threshold();
}

void oracle_world_aabb_f32(const or_shape* s, const float* pos, const float* quat, float* out6) {
  std::vector<V3<float>> st;
  Shape<float> sh = from_desc<float>(s, st);
  Aabb<float> a = sh.world_aabb(v3p<float>(pos), q4p<float>(quat));
  out6[0] = a.min.x; out6[1] = a.min.y; out6[2] = a.min.z;
  out6[3] = a.max.x; out6[4] = a.max.y; out6[5] = a.max.z;
}

void oracle_support_world_f32(const or_shape* s, const float* pos, const float* quat, const float* dir, int core,
                              float* out3) {
  std::vector<V3<float>> st;
  Shape<float> sh = from_desc<float>(s, st);
  V3<float> r = core ? sh.core_support_world(v3p<float>(pos), q4p<float>(quat), v3p<float>(dir))
                     : sh.support_world(v3p<float>(pos), q4p<float>(quat), v3p<float>(dir));
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}

// One collide_seeded call (narrowphase/mod.rs:63-286).  Returns the number
// of contact points (0 = None); out holds 4 x (pos3, normal3, depth).
int oracle_collide_f32(const or_shape* a, const float* pa, const float* qa, const or_shape* b, const float* pb,
                       const float* qb, const float* seed, float* out, int* path) {
  std::vector<V3<float>> sa_, sb_;
  Shape<float> sa = from_desc<float>(a, sa_), sb = from_desc<float>(b, sb_);
  Manifold<float> m;
  V3<float> sd;
  if (seed) sd = v3p<float>(seed);
  int pth = 3;
  collide_seeded(sa, v3p<float>(pa), q4p<float>(qa), sb, v3p<float>(pb), q4p<float>(qb), seed ? &sd : nullptr, m,
                 &pth);
  if (path) *path = pth;
  write_manifold(m, out);
  return m.n;
}

int oracle_collide_f64(const or_shape* a, const double* pa, const double* qa, const or_shape* b, const double* pb,
                       const double* qb, const double* seed, double* out, int* path) {
  std::vector<V3<double>> sa_, sb_;
  Shape<double> sa = from_desc<double>(a, sa_), sb = from_desc<double>(b, sb_);
  Manifold<double> m;
  V3<double> sd;
  if (seed) sd = v3d<double>(seed);
  int pth = 3;
  collide_seeded(sa, v3d<double>(pa), q4d<double>(qa), sb, v3d<double>(pb), q4d<double>(qb), seed ? &sd : nullptr,
                 m, &pth);
  if (path) *path = pth;
  write_manifold(m, out);
  return m.n;
}

// convex_convex (convex_convex.rs:30-155) directly, bypassing the dispatch
// (the reference tests call it on box/sphere pairs too).  `threshold` < 0
// selects the pair's own band (narrowphase/mod.rs:52-57).
int oracle_convex_convex_f32(const or_shape* a, const float* pa, const float* qa, const or_shape* b,
                             const float* pb, const float* qb, const float* seed, float threshold, float* out7,
                             int* path) {
  std::vector<V3<float>> sa_, sb_;
  Shape<float> sa = from_desc<float>(a, sa_), sb = from_desc<float>(b, sb_);
  PairPose<float> pr{&sa, v3p<float>(pa), q4p<float>(qa), &sb, v3p<float>(pb), q4p<float>(qb)};
  V3<float> sd;
  if (seed) sd = v3p<float>(seed);
  float thr = threshold < 0.0f ? pair_threshold(sa, sb) : threshold;
  Manifold<float> m;
  int pth = 0;
  convex_convex(pr, seed ? &sd : nullptr, thr, m, &pth);
  if (path) *path = pth;
  write_manifold(m, out7);
  return m.n;
}

// signed_distance (oracle.hpp) of one pair: out10 = sd, n3, pa3, pb3;
// returns the SdStatus.
extern "C++" template <class T>
int sd_one(const or_shape* a, const T* pa, const T* qa, const or_shape* b, const T* pb, const T* qb,
                  T* out10) {
  std::vector<V3<T>> sa_, sb_;
  Shape<T> sa = from_desc<T>(a, sa_), sb = from_desc<T>(b, sb_);
  PairPose<T> pr{&sa, v3p<T>(pa), q4p<T>(qa), &sb, v3p<T>(pb), q4p<T>(qb)};
  SdOut<T> o = signed_distance(pr);
  T v[10] = {o.sd, o.n.x, o.n.y, o.n.z, o.pa.x, o.pa.y, o.pa.z, o.pb.x, o.pb.y, o.pb.z};
  for (int k = 0; k < 10; ++k) out10[k] = v[k];
  return o.status;
}
int oracle_sd_f32(const or_shape* a, const float* pa, const float* qa, const or_shape* b, const float* pb,
                  const float* qb, float* out10) {
  return sd_one<float>(a, pa, qa, b, pb, qb, out10);
}
int oracle_sd_f64(const or_shape* a, const double* pa, const double* qa, const or_shape* b, const double* pb,
                  const double* qb, double* out10) {
  return sd_one<double>(a, pa, qa, b, pb, qb, out10);
}

// gjk_contact (gjk.rs:171-288): returns 0 Separated / 1 Band / 2 Overlap.
// out7 = distance, normal3, point_on_b3; simplex12 = the Overlap tetrahedron.
int oracle_gjk_contact_f32(const or_shape* a, const float* pa, const float* qa, const or_shape* b, const float* pb,
                           const float* qb, const float* seed, float threshold, float* out7, float* simplex12) {
  std::vector<V3<float>> sa_, sb_;
  Shape<float> sa = from_desc<float>(a, sa_), sb = from_desc<float>(b, sb_);
  PairPose<float> pr{&sa, v3p<float>(pa), q4p<float>(qa), &sb, v3p<float>(pb), q4p<float>(qb)};
  V3<float> sd;
  if (seed) sd = v3p<float>(seed);
  GjkOut<float> g = gjk_contact(pr, seed ? &sd : nullptr, threshold);
  out7[0] = g.distance;
  out7[1] = g.normal.x; out7[2] = g.normal.y; out7[3] = g.normal.z;
  out7[4] = g.point_on_b.x; out7[5] = g.point_on_b.y; out7[6] = g.point_on_b.z;
  if (simplex12)
    for (int i = 0; i < 4; ++i) {
      simplex12[3 * i] = g.simplex[i].x; simplex12[3 * i + 1] = g.simplex[i].y; simplex12[3 * i + 2] = g.simplex[i].z;
    }
  return g.kind;
}

// The walk's exit of gjk_contact in f32: returns the outcome kind and sets
// *iterations and *capped (the 64-iteration cap, reported as separated).
int oracle_gjk_exit_f32(const or_shape* a, const float* pa, const float* qa, const or_shape* b, const float* pb,
                        const float* qb, float threshold, int* iterations, int* capped) {
  std::vector<V3<float>> sa_, sb_;
  Shape<float> sa = from_desc<float>(a, sa_), sb = from_desc<float>(b, sb_);
  PairPose<float> pr{&sa, v3p<float>(pa), q4p<float>(qa), &sb, v3p<float>(pb), q4p<float>(qb)};
  GjkOut<float> g = gjk_contact(pr, (const V3<float>*)nullptr, threshold);
  *iterations = g.iterations;
  *capped = g.capped ? 1 : 0;
  return g.kind;
}

int oracle_gjk_contact_f64(const or_shape* a, const double* pa, const double* qa, const or_shape* b,
                           const double* pb, const double* qb, const double* seed, double threshold, double* out7) {
  std::vector<V3<double>> sa_, sb_;
  Shape<double> sa = from_desc<double>(a, sa_), sb = from_desc<double>(b, sb_);
  PairPose<double> pr{&sa, v3d<double>(pa), q4d<double>(qa), &sb, v3d<double>(pb), q4d<double>(qb)};
  V3<double> sd;
  if (seed) sd = v3d<double>(seed);
  GjkOut<double> g = gjk_contact(pr, seed ? &sd : nullptr, threshold);
  out7[0] = g.distance;
  out7[1] = g.normal.x; out7[2] = g.normal.y; out7[3] = g.normal.z;
  out7[4] = g.point_on_b.x; out7[5] = g.point_on_b.y; out7[6] = g.point_on_b.z;
  return g.kind;
}

// gjk_intersect_seeded (gjk.rs:590-607): 1 when an enclosing simplex exists.
int oracle_gjk_intersect_f32(const or_shape* a, const float* pa, const float* qa, const or_shape* b,
                             const float* pb, const float* qb, float* simplex12) {
  std::vector<V3<float>> sa_, sb_;
  Shape<float> sa = from_desc<float>(a, sa_), sb = from_desc<float>(b, sb_);
  PairPose<float> pr{&sa, v3p<float>(pa), q4p<float>(qa), &sb, v3p<float>(pb), q4p<float>(qb)};
  Simplex<float> s;
  bool ok = gjk_intersect(pr, (const V3<float>*)nullptr, s);
  if (simplex12)
    for (int i = 0; i < 4; ++i) {
      simplex12[3 * i] = s.pts[i].x; simplex12[3 * i + 1] = s.pts[i].y; simplex12[3 * i + 2] = s.pts[i].z;
    }
  return ok ? s.n : 0;
}

// EPA over the full (core=0) or core (core=1) Minkowski support of a pair.
// out4 = normal3, depth.  Returns the EpaStatus (epa.rs:66-83).
int oracle_epa_pair_f32(const float* simplex12, const or_shape* a, const float* pa, const float* qa,
                        const or_shape* b, const float* pb, const float* qb, int core, float* out4) {
  std::vector<V3<float>> sa_, sb_;
  Shape<float> sa = from_desc<float>(a, sa_), sb = from_desc<float>(b, sb_);
  PairPose<float> pr{&sa, v3p<float>(pa), q4p<float>(qa), &sb, v3p<float>(pb), q4p<float>(qb)};
  V3<float> s[4];
  for (int i = 0; i < 4; ++i) s[i] = v3p<float>(simplex12 + 3 * i);
  EpaOut<float> e;
  if (core) e = epa_penetration<float>(s, [&](V3<float> d) { return pr.core_support(d).w; });
  else e = epa_penetration<float>(s, [&](V3<float> d) { return pr.full_support(d); });
  out4[0] = e.normal.x; out4[1] = e.normal.y; out4[2] = e.normal.z; out4[3] = e.depth;
  return e.status;
}

// EPA against a finite point set's support (the reference's point_support
// helper, epa.rs:490-503): first maximum wins.
int oracle_epa_points_f32(const float* simplex12, const float* pts, int npts, float* out4) {
  std::vector<V3<float>> P(npts);
  for (int i = 0; i < npts; ++i) P[i] = v3p<float>(pts + 3 * i);
  V3<float> s[4];
  for (int i = 0; i < 4; ++i) s[i] = v3p<float>(simplex12 + 3 * i);
  auto sup = [&](V3<float> d) {
    V3<float> best = P[0];
    float bd = dot(best, d);
    for (int i = 1; i < npts; ++i) {
      float dd = dot(P[i], d);
      if (dd > bd) { best = P[i]; bd = dd; }
    }
    return best;
  };
  EpaOut<float> e = epa_penetration<float>(s, sup);
  out4[0] = e.normal.x; out4[1] = e.normal.y; out4[2] = e.normal.z; out4[3] = e.depth;
  return e.status;
}

// World AABBs of a scene expanded by `margin` (world/step.rs:349-363).
void oracle_scene_aabbs_f32(const or_scene* sc, float margin, float* out_min, float* out_max) {
  SceneShapes<float> ss;
  ss.build(sc);
  for (int64_t i = 0; i < sc->n; ++i) {
    Aabb<float> a = ss.shapes[i].world_aabb(ss.pos[i], ss.quat[i]).expanded(margin);
    out_min[3 * i] = a.min.x; out_min[3 * i + 1] = a.min.y; out_min[3 * i + 2] = a.min.z;
    out_max[3 * i] = a.max.x; out_max[3 * i + 1] = a.max.y; out_max[3 * i + 2] = a.max.z;
  }
}

// The step's AABB update (step.rs:340-363): predicted positions
// position + linear_velocity * dt for active bodies (flags bit 1; all bodies
// when flags is NULL, body.rs:1175-1181), world AABBs at the predicted pose
// expanded by the margin; inactive bodies keep prev_min / prev_max when
// given.  out_pos receives the predicted positions.
void oracle_scene_aabbs_predicted_f32(const or_scene* sc, const float* vel, float dt, const uint8_t* flags,
                                      const float* prev_min, const float* prev_max, float margin, float* out_pos,
                                      float* out_min, float* out_max) {
  SceneShapes<float> ss;
  ss.build(sc);
  for (int64_t i = 0; i < sc->n; ++i) {
    bool active = !flags || (flags[i] & 2u);
    V3<float> p = ss.pos[i];
    if (active) p = p + mk(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]) * dt;
    out_pos[3 * i] = p.x; out_pos[3 * i + 1] = p.y; out_pos[3 * i + 2] = p.z;
    if (!active && prev_min && prev_max) {
      for (int k = 0; k < 3; ++k) {
        out_min[3 * i + k] = prev_min[3 * i + k];
        out_max[3 * i + k] = prev_max[3 * i + k];
      }
      continue;
    }
    Aabb<float> a = ss.shapes[i].world_aabb(p, ss.quat[i]).expanded(margin);
    out_min[3 * i] = a.min.x; out_min[3 * i + 1] = a.min.y; out_min[3 * i + 2] = a.min.z;
    out_max[3 * i] = a.max.x; out_max[3 * i + 1] = a.max.y; out_max[3 * i + 2] = a.max.z;
  }
}

// Per-body contact_breaking_threshold (shapes/mod.rs:388-392).
void oracle_scene_thresholds_f32(const or_scene* sc, float* out) {
  SceneShapes<float> ss;
  ss.build(sc);
  for (int64_t i = 0; i < sc->n; ++i) out[i] = ss.shapes[i].contact_breaking_threshold();
}

// BruteForceBroadPhase::compute_pairs_into (brute_force.rs:54-67) over
// AABBs given as min/max triplets.  Writes up to cap pairs, returns the total.
int64_t oracle_brute_pairs(int64_t n, const float* mn, const float* mx, uint32_t* out, int64_t cap) {
  int64_t cnt = 0;
  for (int64_t i = 0; i < n; ++i) {
    Aabb<float> a{v3p<float>(mn + 3 * i), v3p<float>(mx + 3 * i)};
    for (int64_t j = i + 1; j < n; ++j) {
      Aabb<float> b{v3p<float>(mn + 3 * j), v3p<float>(mx + 3 * j)};
      if (a.overlaps(b)) {
        if (cnt < cap) { out[2 * cnt] = (uint32_t)i; out[2 * cnt + 1] = (uint32_t)j; }
        ++cnt;
      }
    }
  }
  return cnt;
}

// aabb_raycast_candidates (broadphase/mod.rs:185-200) over n boxes.
int64_t oracle_raycast_candidates(int64_t n, const float* mn, const float* mx, const float* from, const float* to,
                                  uint32_t* out, int64_t cap) {
  int64_t cnt = 0;
  for (int64_t i = 0; i < n; ++i) {
    Aabb<float> a{v3p<float>(mn + 3 * i), v3p<float>(mx + 3 * i)};
    if (segment_hits_aabb(a, v3p<float>(from), v3p<float>(to))) {
      if (cnt < cap) out[cnt] = (uint32_t)i;
      ++cnt;
    }
  }
  return cnt;
}

// The same pair set through an x-axis sweep (for sizes brute force cannot
// reach); checked against oracle_brute_pairs by the tests.
int64_t oracle_sweep_pairs(int64_t n, const float* mn, const float* mx, uint32_t* out, int64_t cap) {
  std::vector<Aabb<float>> boxes(n);
  for (int64_t i = 0; i < n; ++i) boxes[i] = Aabb<float>{v3p<float>(mn + 3 * i), v3p<float>(mx + 3 * i)};
  std::vector<uint64_t> keys;
  int64_t cnt = sweep_pairs<float>(n, boxes.data(), keys);
  for (int64_t k = 0; k < cnt && k < cap; ++k) {
    out[2 * k] = (uint32_t)(keys[k] >> 32);
    out[2 * k + 1] = (uint32_t)(keys[k] & 0xffffffffu);
  }
  return cnt;
}

// The same pair set through the DbvtBroadPhase restatement (dbvt_pairs);
// checked against oracle_brute_pairs by the tests.
int64_t oracle_dbvt_pairs(int64_t n, const float* mn, const float* mx, uint32_t* out, int64_t cap) {
  std::vector<Aabb<float>> boxes(n);
  for (int64_t i = 0; i < n; ++i) boxes[i] = Aabb<float>{v3p<float>(mn + 3 * i), v3p<float>(mx + 3 * i)};
  std::vector<uint64_t> keys;
  int64_t cnt = dbvt_pairs<float>(n, boxes.data(), keys);
  for (int64_t k = 0; k < cnt && k < cap; ++k) {
    out[2 * k] = (uint32_t)(keys[k] >> 32);
    out[2 * k + 1] = (uint32_t)(keys[k] & 0xffffffffu);
  }
  return cnt;
}

// collide_seeded for each (i, j) pair of a scene (shape_a = body i).
// out_count[P] (0..4), out_pts[P*28] (4 x pos3 normal3 depth), out_path[P]
// (0 gjk-separated, 1 gjk-band, 2 epa, 3 closed form).
void oracle_collide_pairs_f32(const or_scene* sc, const uint32_t* pairs, int64_t npairs, const float* seeds,
                              uint8_t* out_count, float* out_pts, int8_t* out_path) {
  SceneShapes<float> ss;
  ss.build(sc);
  for (int64_t k = 0; k < npairs; ++k) {
    uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    Manifold<float> m;
    V3<float> sd;
    bool has_seed = seeds && !std::isnan(seeds[3 * k]);  // NaN x: None for this pair
    if (has_seed) sd = v3p<float>(seeds + 3 * k);
    int pth = 3;
    collide_seeded(ss.shapes[i], ss.pos[i], ss.quat[i], ss.shapes[j], ss.pos[j], ss.quat[j], has_seed ? &sd : nullptr,
                   m, &pth);
    out_count[k] = (uint8_t)m.n;
    if (out_path) out_path[k] = (int8_t)pth;
    if (out_pts) write_manifold(m, out_pts + 28 * k);
  }
}

// signed_distance over scene pairs.  pose14 (per pair: pos_a3 quat_a4 pos_b3
// quat_b4) overrides the scene poses when given (finite differences).
extern "C++" template <class T, class P>
void sd_pairs(const or_scene* sc, const uint32_t* pairs, int64_t npairs, const P* pose14, T* out10,
                     int8_t* status) {
  SceneShapes<T> ss;
  ss.build(sc);
  for (int64_t k = 0; k < npairs; ++k) {
    uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    PairPose<T> pr{&ss.shapes[i], ss.pos[i], ss.quat[i], &ss.shapes[j], ss.pos[j], ss.quat[j]};
    if (pose14) {
      const P* q = pose14 + 14 * k;
      pr.pa = v3p<T>(q);
      pr.qa = q4p<T>(q + 3);
      pr.pb = v3p<T>(q + 7);
      pr.qb = q4p<T>(q + 10);
    }
    SdOut<T> o = signed_distance(pr);
    T v[10] = {o.sd, o.n.x, o.n.y, o.n.z, o.pa.x, o.pa.y, o.pa.z, o.pb.x, o.pb.y, o.pb.z};
    for (int c = 0; c < 10; ++c) out10[10 * k + c] = v[c];
    if (status) status[k] = (int8_t)o.status;
  }
}
// The walk's active simplex (Band or capped) of one scene pair: returns its
// size k (0 when the walk did not end in that arm) and w[3k], b[3k].
extern "C++" template <class T>
int sd_active(const or_scene* sc, uint32_t i, uint32_t j, double* w, double* b) {
  SceneShapes<T> ss;
  ss.build(sc);
  PairPose<T> pr{&ss.shapes[i], ss.pos[i], ss.quat[i], &ss.shapes[j], ss.pos[j], ss.quat[j]};
  GjkOut<T> g = gjk_contact(pr, (const V3<T>*)nullptr, std::numeric_limits<T>::infinity(), true);
  if (!(g.kind == 1 || (g.kind == 0 && g.capped))) return 0;
  for (int k = 0; k < g.n_active; ++k) {
    w[3 * k] = g.active[k].w.x; w[3 * k + 1] = g.active[k].w.y; w[3 * k + 2] = g.active[k].w.z;
    b[3 * k] = g.active[k].p_b.x; b[3 * k + 1] = g.active[k].p_b.y; b[3 * k + 2] = g.active[k].p_b.z;
  }
  return g.n_active;
}
int oracle_sd_active(const or_scene* sc, uint32_t i, uint32_t j, int f64, double* w, double* b) {
  return f64 ? sd_active<double>(sc, i, j, w, b) : sd_active<float>(sc, i, j, w, b);
}

// sd_vjp over scene pairs at the scene poses (f64): g10 [10P] -> g14 [14P];
// returns the number of pairs whose EPA face could not be split.
int64_t oracle_sd_vjp_pairs_f64(const or_scene* sc, const uint32_t* pairs, int64_t npairs, const double* g10,
                                double* g14) {
  SceneShapes<double> ss;
  ss.build(sc);
  int64_t fails = 0;
  for (int64_t k = 0; k < npairs; ++k) {
    uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    PairPose<double> pr{&ss.shapes[i], ss.pos[i], ss.quat[i], &ss.shapes[j], ss.pos[j], ss.quat[j]};
    if (!sd_vjp(pr, g10 + 10 * k, g14 + 14 * k)) ++fails;
  }
  return fails;
}

void oracle_sd_pairs_f32(const or_scene* sc, const uint32_t* pairs, int64_t npairs, float* out10, int8_t* status) {
  sd_pairs<float, float>(sc, pairs, npairs, (const float*)nullptr, out10, status);
}
void oracle_sd_pairs_f64(const or_scene* sc, const uint32_t* pairs, int64_t npairs, const double* pose14,
                         double* out10, int8_t* status) {
  sd_pairs<double, double>(sc, pairs, npairs, pose14, out10, status);
}

void oracle_collide_pairs_f64(const or_scene* sc, const uint32_t* pairs, int64_t npairs, uint8_t* out_count,
                              double* out_pts) {
  SceneShapes<double> ss;
  ss.build(sc);
  for (int64_t k = 0; k < npairs; ++k) {
    uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    Manifold<double> m;
    collide_seeded(ss.shapes[i], ss.pos[i], ss.quat[i], ss.shapes[j], ss.pos[j], ss.quat[j], (const V3<double>*)nullptr, m);
    out_count[k] = (uint8_t)m.n;
    write_manifold(m, out_pts + 28 * k);
  }
}

// One full collision pass on the CPU, the timed CPU baseline: world AABBs +
// margin, the exact pair set (sweep), collide_seeded(body i, body j, None)
// for every pair.  Returns the pair count; *contacts = total contact points;
// *seconds = wall time of the pass.
int64_t oracle_pass_f32(const or_scene* sc, float margin, int64_t* contacts, double* seconds) {
  SceneShapes<float> ss;
  ss.build(sc);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<Aabb<float>> boxes(sc->n);
  for (int64_t i = 0; i < sc->n; ++i) boxes[i] = ss.shapes[i].world_aabb(ss.pos[i], ss.quat[i]).expanded(margin);
  std::vector<uint64_t> keys;
  int64_t np = dbvt_pairs<float>(sc->n, boxes.data(), keys);
  int64_t nc = 0;
  for (int64_t k = 0; k < np; ++k) {
    uint32_t i = (uint32_t)(keys[k] >> 32), j = (uint32_t)(keys[k] & 0xffffffffu);
    Manifold<float> m;
    collide_seeded(ss.shapes[i], ss.pos[i], ss.quat[i], ss.shapes[j], ss.pos[j], ss.quat[j],
                   (const V3<float>*)nullptr, m);
    nc += m.n;
  }
  auto t1 = std::chrono::steady_clock::now();
  *contacts = nc;
  *seconds = std::chrono::duration<double>(t1 - t0).count();
  return np;
}

// The same pass with the per-body AABBs and the per-pair narrowphase spread
// over `threads` OpenMP threads (the pair sweep stays serial): the all-cores
// CPU figure.  Pairs are independent, so the counts equal the serial pass's.
int64_t oracle_pass_f32_mt(const or_scene* sc, float margin, int threads, int64_t* contacts, double* seconds) {
  SceneShapes<float> ss;
  ss.build(sc);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<Aabb<float>> boxes(sc->n);
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int64_t i = 0; i < sc->n; ++i) boxes[i] = ss.shapes[i].world_aabb(ss.pos[i], ss.quat[i]).expanded(margin);
  std::vector<uint64_t> keys;
  int64_t np = dbvt_pairs<float>(sc->n, boxes.data(), keys);
  int64_t nc = 0;
#pragma omp parallel for schedule(dynamic, 512) reduction(+ : nc) num_threads(threads)
  for (int64_t k = 0; k < np; ++k) {
    uint32_t i = (uint32_t)(keys[k] >> 32), j = (uint32_t)(keys[k] & 0xffffffffu);
    Manifold<float> m;
    collide_seeded(ss.shapes[i], ss.pos[i], ss.quat[i], ss.shapes[j], ss.pos[j], ss.quat[j],
                   (const V3<float>*)nullptr, m);
    nc += m.n;
  }
  auto t1 = std::chrono::steady_clock::now();
  *contacts = nc;
  *seconds = std::chrono::duration<double>(t1 - t0).count();
  return np;
}

// ---------------------------------------------------- persistent manifold --
// PersistentManifold operations on one serialised entry (64 floats, layout in
// oracle.hpp::pm_store).  Poses are (x, y, z) and (x, y, z, w).
static void pm_emit(const PContact& c, V3<float> anchor, float* o) {
  o[0] = c.position.x; o[1] = c.position.y; o[2] = c.position.z;
  o[3] = c.normal.x; o[4] = c.normal.y; o[5] = c.normal.z;
  o[6] = c.penetration; o[7] = c.normal_impulse;
  std::memcpy(o + 8, &c.lifetime, 4);
  o[9] = anchor.x; o[10] = anchor.y; o[11] = anchor.z;
}

// capture (manifold.rs:320-376) of n points (7 floats: position, normal, penetration).
void oracle_pm_capture(float* e64, const float* pts7, int n, const float* pa, const float* qa, const float* pb,
                       const float* qb, float threshold) {
  PMan m;
  pm_load(e64, m);
  ContactPoint<float> cps[8];
  for (int k = 0; k < n && k < 8; ++k)
    cps[k] = ContactPoint<float>{v3p<float>(pts7 + 7 * k), v3p<float>(pts7 + 7 * k + 3), pts7[7 * k + 6]};
  pm_capture(m, cps, n < 8 ? n : 8, v3p<float>(pa), q4p<float>(qa), v3p<float>(pb), q4p<float>(qb), threshold);
  pm_store(m, e64);
}

// refresh (manifold.rs:252-298): returns the surviving count; out12[k] as
// the pass emits it (impulse unscaled, anchor = pa + qa * local_a).
int oracle_pm_refresh(float* e64, const float* pa, const float* qa, const float* pb, const float* qb, float* out12) {
  PMan m;
  pm_load(e64, m);
  PContact out[PM_CACHE];
  int n = pm_refresh(m, v3p<float>(pa), q4p<float>(qa), v3p<float>(pb), q4p<float>(qb), out);
  for (int k = 0; k < n; ++k)
    pm_emit(out[k], v3p<float>(pa) + qrot(q4p<float>(qa), m.pts[k].local_a), out12 + 12 * k);
  pm_store(m, e64);
  return n;
}

int oracle_pm_needs_redetect(const float* e64, const float* pa, const float* qa, const float* pb, const float* qb) {
  PMan m;
  pm_load(e64, m);
  return pm_needs_redetect(m, v3p<float>(pa), q4p<float>(qa), v3p<float>(pb), q4p<float>(qb)) ? 1 : 0;
}

int oracle_pm_seed(const float* e64, float* out3) {
  PMan m;
  pm_load(e64, m);
  V3<float> s;
  if (!pm_seed(m, &s)) return 0;
  out3[0] = s.x; out3[1] = s.y; out3[2] = s.z;
  return 1;
}

// One step of the pair pass with persistent manifolds (step.rs:660-854) over
// the sorted pairs of a scene.  The cache comes in as the previous step's
// sorted keys with entries and valid flags, and goes out per pair of this
// step (entries of pairs absent from this step are pruned, step.rs:1006-1010).
// flags (NULL: every body active, dynamic, slow) and group_mask (group | mask
// << 16; NULL: everything collides) per body.  Per pair: route, entry valid,
// entry, emitted count and 4 x 12 floats (position, normal, penetration,
// impulse x warm, lifetime bits, A-side anchor).
void oracle_pm_pass(const or_scene* sc, const uint32_t* pairs, int64_t npairs, const uint8_t* flags,
                    const uint32_t* group_mask, float slop, float warm, const uint64_t* keys_in, const float* ent_in,
                    const uint8_t* valid_in, int64_t n_in, uint8_t* route, uint8_t* valid_out, float* ent_out,
                    uint8_t* out_cnt, float* out_pts) {
  SceneShapes<float> ss;
  ss.build(sc);
  for (int64_t p = 0; p < npairs; ++p) {
    uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
    uint64_t key = ((uint64_t)i << 32) | j;
    const uint64_t* it = std::lower_bound(keys_in, keys_in + n_in, key);
    int64_t e = (it != keys_in + n_in && *it == key) ? (int64_t)(it - keys_in) : -1;
    bool has = e >= 0 && valid_in[e];
    PMan m;
    if (has) pm_load(ent_in + 64 * e, m);
    uint8_t fa = flags ? flags[i] : (uint8_t)PM_ACTIVE, fb = flags ? flags[j] : (uint8_t)PM_ACTIVE;
    uint32_t ga = group_mask ? group_mask[i] : 0xffffffffu, gb = group_mask ? group_mask[j] : 0xffffffffu;
    PmPassOut r = pm_pair_pass(has, m, ss.shapes[i], ss.pos[i], ss.quat[i], ss.shapes[j], ss.pos[j], ss.quat[j], fa, fb,
                               ga, gb, slop, warm);
    route[p] = r.route;
    valid_out[p] = r.keep ? 1 : 0;
    if (r.keep) pm_store(m, ent_out + 64 * p);
    else std::memset(ent_out + 64 * p, 0, 64 * sizeof(float));
    out_cnt[p] = (uint8_t)r.n_out;
    std::memset(out_pts + 48 * p, 0, 48 * sizeof(float));
    for (int k = 0; k < r.n_out; ++k) pm_emit(r.out[k], r.anchor_a[k], out_pts + 48 * p + 12 * k);
  }
}

// ----------------------------------------------------------- triangle meshes --
void* oracle_mesh_create(const float* verts, uint32_t nv, const uint32_t* ind, uint32_t nt) {
  TriMesh* m = new TriMesh();
  m->build(verts, nv, ind, nt);
  return m;
}

void oracle_mesh_destroy(void* h) { delete static_cast<TriMesh*>(h); }

float oracle_mesh_threshold(const void* h) { return static_cast<const TriMesh*>(h)->threshold(); }

// triangles_in_aabb (triangle_mesh.rs:150-154, bvh.rs:58-79) in query order.
int64_t oracle_mesh_query(const void* h, const float* qmin, const float* qmax, uint32_t* out, int64_t cap) {
  std::vector<uint32_t> c;
  static_cast<const TriMesh*>(h)->query(Aabb<float>{v3p<float>(qmin), v3p<float>(qmax)}, c);
  for (size_t k = 0; k < c.size() && (int64_t)k < cap; ++k) out[k] = c[k];
  return (int64_t)c.size();
}

// collide_seeded's convex-vs-mesh arms (narrowphase/mod.rs:177-215): the
// pair threshold min(convex, mesh) (mod.rs:52-57); with mesh_first the mesh
// is shape A and the manifold is flipped.  Returns the point count; out 4 x 7.
static int mesh_pair(const Shape<float>& convex, V3<float> pc, Quat<float> qc, const TriMesh& mesh, V3<float> pm,
                     Quat<float> qm, const float* seed, int mesh_first, float* out) {
  float thr = fmin_(convex.contact_breaking_threshold(), mesh.threshold());
  if (mesh_first) thr = fmin_(mesh.threshold(), convex.contact_breaking_threshold());
  V3<float> sd;
  bool has_seed = seed && !std::isnan(seed[0]);
  if (has_seed) sd = v3p<float>(seed);
  std::vector<ContactPoint<float>> pts;
  if (!convex_mesh(convex, pc, qc, mesh, pm, qm, has_seed ? &sd : nullptr, thr, pts)) return 0;
  for (size_t k = 0; k < pts.size() && k < 4; ++k) {
    float* o = out + 7 * k;
    V3<float> n = mesh_first ? -pts[k].normal : pts[k].normal;
    o[0] = pts[k].position.x; o[1] = pts[k].position.y; o[2] = pts[k].position.z;
    o[3] = n.x; o[4] = n.y; o[5] = n.z;
    o[6] = pts[k].penetration;
  }
  return (int)pts.size();
}

int oracle_mesh_collide(const or_shape* convex, const float* pc, const float* qc, const void* mesh, const float* pm,
                        const float* qm, const float* seed, int mesh_first, float* out) {
  std::vector<V3<float>> st;
  Shape<float> cv = from_desc<float>(convex, st);
  return mesh_pair(cv, v3p<float>(pc), q4p<float>(qc), *static_cast<const TriMesh*>(mesh), v3p<float>(pm),
                   q4p<float>(qm), seed, mesh_first, out);
}

// A batch over scene bodies: pair k = (body pairs[2k], mesh handle
// meshes[pairs[2k+1]]) at mesh pose mesh_pose[7k..] (position, orientation);
// seeds[3k..] (NULL, or NaN x = None); mesh_first[k] (NULL = 0).
void oracle_mesh_collide_scene(const or_scene* sc, void* const* meshes, const uint32_t* pairs, int64_t npairs,
                               const float* mesh_pose, const float* seeds, const uint8_t* mesh_first,
                               uint8_t* out_count, float* out_pts) {
  SceneShapes<float> ss;
  ss.build(sc);
  for (int64_t k = 0; k < npairs; ++k) {
    uint32_t b = pairs[2 * k];
    const TriMesh& m = *static_cast<const TriMesh*>(meshes[pairs[2 * k + 1]]);
    std::memset(out_pts + 28 * k, 0, 28 * sizeof(float));
    out_count[k] = (uint8_t)mesh_pair(ss.shapes[b], ss.pos[b], ss.quat[b], m, v3p<float>(mesh_pose + 7 * k),
                                      q4p<float>(mesh_pose + 7 * k + 3), seeds ? seeds + 3 * k : nullptr,
                                      mesh_first ? mesh_first[k] : 0, out_pts + 28 * k);
  }
}

// sphere_triangle_contact (convex_mesh.rs:168-235): 1 and point, normal,
// penetration when it reports a contact.
int oracle_sphere_triangle(const float* center, float radius, float threshold, const float* a, const float* b,
                           const float* c, float* out7) {
  V3<float> p, n;
  float pen;
  if (!sphere_triangle_contact(v3p<float>(center), radius, threshold, v3p<float>(a), v3p<float>(b), v3p<float>(c), p, n,
                               pen))
    return 0;
  out7[0] = p.x; out7[1] = p.y; out7[2] = p.z;
  out7[3] = n.x; out7[4] = n.y; out7[5] = n.z;
  out7[6] = pen;
  return 1;
}

// ContactManifold::reduce (manifold.rs:76-110) of n points (7 floats each) to
// at most max, in place; returns the new count.
int oracle_reduce(float* pts7, int n, int max) {
  std::vector<ContactPoint<float>> v(n);
  for (int k = 0; k < n; ++k)
    v[k] = ContactPoint<float>{v3p<float>(pts7 + 7 * k), v3p<float>(pts7 + 7 * k + 3), pts7[7 * k + 6]};
  reduce_manifold(v, (size_t)max);
  for (size_t k = 0; k < v.size(); ++k) {
    float* o = pts7 + 7 * k;
    o[0] = v[k].position.x; o[1] = v[k].position.y; o[2] = v[k].position.z;
    o[3] = v[k].normal.x; o[4] = v[k].normal.y; o[5] = v[k].normal.z;
    o[6] = v[k].penetration;
  }
  return (int)v.size();
}

// ---------------------------------------------------------------- sweeps --
static int emit_hit(bool hit, const SweepHit& h, float* out7) {
  if (!hit) {
    for (int k = 0; k < 7; ++k) out7[k] = 0.0f;
    return 0;
  }
  out7[0] = h.fraction;
  out7[1] = h.normal.x; out7[2] = h.normal.y; out7[3] = h.normal.z;
  out7[4] = h.point.x; out7[5] = h.point.y; out7[6] = h.point.z;
  return 1;
}

// swept_sphere_cast against a convex / plane shape: out7 = fraction, normal, point.
int oracle_sweep_sphere(const or_shape* other, const float* op, const float* oq, const float* from, const float* to,
                        float radius, float* out7) {
  std::vector<V3<float>> st;
  Shape<float> o = from_desc<float>(other, st);
  SweepHit h;
  bool hit = sweep_sphere(v3p<float>(from), v3p<float>(to), radius, o, v3p<float>(op), q4p<float>(oq), h);
  return emit_hit(hit, h, out7);
}

int oracle_sweep_sphere_mesh(const void* mesh, const float* pos, const float* ori, const float* from, const float* to,
                             float radius, float* out7) {
  SweepHit h;
  bool hit = sweep_sphere_mesh(v3p<float>(from), v3p<float>(to), radius, *static_cast<const TriMesh*>(mesh),
                               v3p<float>(pos), q4p<float>(ori), h);
  return emit_hit(hit, h, out7);
}

int oracle_convex_cast(const or_shape* shape, const float* from, const float* ori, const float* to,
                       const or_shape* other, const float* op, const float* oq, float* out7) {
  std::vector<V3<float>> s1, s2;
  Shape<float> a = from_desc<float>(shape, s1), o = from_desc<float>(other, s2);
  SweepHit h;
  bool hit = convex_cast(a, v3p<float>(from), q4p<float>(ori), v3p<float>(to), o, v3p<float>(op), q4p<float>(oq), h);
  return emit_hit(hit, h, out7);
}

// Batches over scene bodies (targets at their scene poses).  Sphere sweeps:
// query k casts radius[k] from from[3k] to to[3k] against body target[k].
// Convex casts: body caster[k]'s shape at its scene orientation from from[3k]
// to to[3k] against body target[k].  hit[k], out7[7k ..].
void oracle_sweep_scene(const or_scene* sc, const uint32_t* target, const float* from, const float* to,
                        const float* radius, const uint32_t* caster, int64_t n, uint8_t* hit, float* out7) {
  SceneShapes<float> ss;
  ss.build(sc);
  for (int64_t k = 0; k < n; ++k) {
    uint32_t t = target[k];
    SweepHit h;
    bool ok;
    if (caster) {
      uint32_t c = caster[k];
      ok = convex_cast(ss.shapes[c], v3p<float>(from + 3 * k), ss.quat[c], v3p<float>(to + 3 * k), ss.shapes[t],
                       ss.pos[t], ss.quat[t], h);
    } else {
      ok = sweep_sphere(v3p<float>(from + 3 * k), v3p<float>(to + 3 * k), radius[k], ss.shapes[t], ss.pos[t],
                        ss.quat[t], h);
    }
    hit[k] = (uint8_t)emit_hit(ok, h, out7 + 7 * k);
  }
}

// Sweeps against meshes over scene casters: query k casts (sphere radius[k]
// when caster is NULL, else body caster[k]'s shape at its scene orientation)
// from from[3k] to to[3k] against mesh meshes[mesh_of[k]] at mesh_pose[7k ..].
void oracle_sweep_mesh_scene(const or_scene* sc, void* const* meshes, const uint32_t* mesh_of, const float* mesh_pose,
                             const float* from, const float* to, const float* radius, const uint32_t* caster,
                             int64_t n, uint8_t* hit, float* out7) {
  SceneShapes<float> ss;
  if (sc) ss.build(sc);
  for (int64_t k = 0; k < n; ++k) {
    const TriMesh& m = *static_cast<const TriMesh*>(meshes[mesh_of[k]]);
    V3<float> mp = v3p<float>(mesh_pose + 7 * k);
    Quat<float> mq = q4p<float>(mesh_pose + 7 * k + 3);
    SweepHit h;
    bool ok;
    if (caster) {
      uint32_t c = caster[k];
      ok = convex_cast_mesh(ss.shapes[c], v3p<float>(from + 3 * k), ss.quat[c], v3p<float>(to + 3 * k), m, mp, mq, h);
    } else {
      ok = sweep_sphere_mesh(v3p<float>(from + 3 * k), v3p<float>(to + 3 * k), radius[k], m, mp, mq, h);
    }
    hit[k] = (uint8_t)emit_hit(ok, h, out7 + 7 * k);
  }
}

// ------------------------------------------------------------ helpers --
// Helper-level entry points for the known-answer tests of the reference's
// own helper tests (gjk.rs:1233-1537, :1703-2093; epa.rs:480-738) and of the
// tie rules the reference documents (epa.rs:212-224, :333-345;
// box_box.rs:126-174).  All f32.

// closest_origin (gjk.rs:864-1043) on a simplex of n points; returns the
// number of kept points (written to keep12 in the reference's keep order).
int oracle_closest_origin_f32(const float* pts, int n, float* q3, float* keep12) {
  CoSimplex s;
  s.n = n;
  for (int i = 0; i < n; ++i) s.p[i] = v3p<float>(pts + 3 * i);
  V3<float> q = closest_origin(s);
  q3[0] = q.x; q3[1] = q.y; q3[2] = q.z;
  for (int i = 0; i < s.n; ++i) { keep12[3 * i] = s.p[i].x; keep12[3 * i + 1] = s.p[i].y; keep12[3 * i + 2] = s.p[i].z; }
  return s.n;
}

// The contact walk's reduce (gjk.rs:352-583) on n support vertices (w, p_b);
// returns 1 for Inside, else writes v, p_b and the shrunk length.
int oracle_reduce_f32(const float* w12, const float* pb12, int n, float* v3, float* pb3, int* n_out) {
  SupportVertex<float> buf[4];
  for (int i = 0; i < 4; ++i) {
    buf[i].w = i < n ? v3p<float>(w12 + 3 * i) : splat(0.0f);
    buf[i].p_b = i < n ? v3p<float>(pb12 + 3 * i) : splat(0.0f);
  }
  int len = n;
  Reduce<float> r = reduce(buf, len);
  *n_out = len;
  if (r.inside) return 1;
  v3[0] = r.v.x; v3[1] = r.v.y; v3[2] = r.v.z;
  pb3[0] = r.p_b.x; pb3[1] = r.p_b.y; pb3[2] = r.p_b.z;
  return 0;
}

// perpendicular (gjk.rs:668-680)
void oracle_perpendicular_f32(const float* v, float* out3) {
  V3<float> p = perpendicular(v3p<float>(v));
  out3[0] = p.x; out3[1] = p.y; out3[2] = p.z;
}

// solve (gjk.rs:609-620) on the strict-walk simplex built as
// Simplex::new(pts[0]) followed by add(pts[1]), add(pts[2]), ...
int oracle_solve_f32(const float* pts, int n, float* dir3, int* n_out) {
  Simplex<float> s;
  s.init(v3p<float>(pts));
  for (int i = 1; i < n; ++i) s.add(v3p<float>(pts + 3 * i));
  V3<float> dir = splat(0.0f);
  bool r = solve(s, dir);
  dir3[0] = dir.x; dir3[1] = dir.y; dir3[2] = dir.z;
  *n_out = s.n;
  return r ? 1 : 0;
}

// enclosing_walk (gjk.rs:299-342) with a scripted support: call k returns
// script[min(k, nscript - 1)]; every queried direction is recorded (up to
// max_q).  Returns 1 when an enclosing simplex was found.
int oracle_enclosing_walk_scripted_f32(const float* script, int nscript, const float* fallback, float* queried,
                                       int max_q, int* n_queried) {
  int calls = 0;
  auto sup = [&](V3<float> d) {
    if (calls < max_q) { queried[3 * calls] = d.x; queried[3 * calls + 1] = d.y; queried[3 * calls + 2] = d.z; }
    int k = calls < nscript ? calls : nscript - 1;
    ++calls;
    return v3p<float>(script + 3 * k);
  };
  Simplex<float> s;
  bool ok = enclosing_walk<float>(sup, v3p<float>(fallback), (const V3<float>*)nullptr, s);
  *n_queried = calls;
  return ok ? 1 : 0;
}

namespace {
Face<float> make_face(const std::vector<V3<float>>& v, int a, int b, int c) {
  V3<float> cr = cross(v[b] - v[a], v[c] - v[a]);
  V3<float> n = cr / len(cr);
  return Face<float>{a, b, c, n, face_distance(n, v[a], v[b], v[c]), dot(n, v[a])};
}
}  // namespace

// find_best (epa.rs:212-224) over faces whose |distance| values are given.
int oracle_epa_find_best_f32(const float* distance, int n) {
  std::vector<Face<float>> f(n);
  for (int i = 0; i < n; ++i) f[i] = Face<float>{0, 0, 0, splat(0.0f), distance[i], distance[i]};
  return find_best(f);
}

// expand (epa.rs:309-388) of the polytope with faces abc (index triples
// into verts, normals built as the stitch builds them) toward verts[nv - 1].
// Writes the resulting face triples in list order; returns 1 when accepted,
// 0 when the horizon was rejected (faces unchanged).
int oracle_epa_expand_f32(const int* abc, int nf, const float* verts, int nv, int* out_abc, int* out_nf) {
  std::vector<V3<float>> v(nv);
  for (int i = 0; i < nv; ++i) v[i] = v3p<float>(verts + 3 * i);
  std::vector<Face<float>> faces;
  for (int i = 0; i < nf; ++i) faces.push_back(make_face(v, abc[3 * i], abc[3 * i + 1], abc[3 * i + 2]));
  std::vector<std::pair<int, int>> boundary;
  bool ok = expand(faces, v, boundary, v[nv - 1], nv - 1);
  for (size_t i = 0; i < faces.size(); ++i) {
    out_abc[3 * i] = faces[i].a; out_abc[3 * i + 1] = faces[i].b; out_abc[3 * i + 2] = faces[i].c;
  }
  *out_nf = (int)faces.size();
  return ok ? 1 : 0;
}

// cull_points (box_box.rs:131-174): n <= 8 points with depths, keep maxc.
int oracle_cull_points_f32(const float* pts, const float* deps, int n, int maxc, float* out_pts, float* out_deps) {
  V3<float> p[8], op[4];
  float d[8], od[4];
  for (int i = 0; i < n; ++i) { p[i] = v3p<float>(pts + 3 * i); d[i] = deps[i]; }
  int m = cull_points(p, d, n, maxc, op, od);
  for (int i = 0; i < m; ++i) {
    out_pts[3 * i] = op[i].x; out_pts[3 * i + 1] = op[i].y; out_pts[3 * i + 2] = op[i].z;
    out_deps[i] = od[i];
  }
  return m;
}

// EPA (epa.rs:109-205) against a finite point set's support (first maximum
// wins, as the reference's point_support helper), recording every support
// query direction (up to max_q); also reports the status, the number of
// polytope vertices and the main-loop iteration count at exit.
int oracle_epa_points_trace_f32(const float* simplex12, const float* pts, int npts, float* out4, float* queries,
                                int max_q, int* n_queries, int* n_vertices, int* iterations) {
  std::vector<V3<float>> P(npts);
  for (int i = 0; i < npts; ++i) P[i] = v3p<float>(pts + 3 * i);
  V3<float> s[4];
  for (int i = 0; i < 4; ++i) s[i] = v3p<float>(simplex12 + 3 * i);
  int calls = 0;
  auto sup = [&](V3<float> d) {
    if (calls < max_q) { queries[3 * calls] = d.x; queries[3 * calls + 1] = d.y; queries[3 * calls + 2] = d.z; }
    ++calls;
    V3<float> best = P[0];
    float bd = dot(best, d);
    for (int i = 1; i < npts; ++i) {
      float dd = dot(P[i], d);
      if (dd > bd) { best = P[i]; bd = dd; }
    }
    return best;
  };
  EpaOut<float> e = epa_penetration<float>(s, sup);
  out4[0] = e.normal.x; out4[1] = e.normal.y; out4[2] = e.normal.z; out4[3] = e.depth;
  *n_queries = calls;
  *n_vertices = e.vertices;
  *iterations = e.iterations;
  return e.status;
}

// face_distance (epa.rs:262-303) of the triangle a, b, c with the normal
// the polytope gives it (normalized cross(b - a, c - a)).
float oracle_epa_face_distance_f32(const float* a, const float* b, const float* c) {
  V3<float> pa = v3p<float>(a), pb = v3p<float>(b), pc = v3p<float>(c);
  V3<float> cr = cross(pb - pa, pc - pa);
  return face_distance(cr / len(cr), pa, pb, pc);
}

// oracle_epa_pair_f32 that also reports the polytope's vertex count and the
// main-loop iteration count at exit.
int oracle_epa_pair_trace_f32(const float* simplex12, const or_shape* a, const float* pa, const float* qa,
                              const or_shape* b, const float* pb, const float* qb, int core, float* out4,
                              int* n_vertices, int* iterations) {
  std::vector<V3<float>> sa_, sb_;
  Shape<float> sa = from_desc<float>(a, sa_), sb = from_desc<float>(b, sb_);
  PairPose<float> pr{&sa, v3p<float>(pa), q4p<float>(qa), &sb, v3p<float>(pb), q4p<float>(qb)};
  V3<float> s[4];
  for (int i = 0; i < 4; ++i) s[i] = v3p<float>(simplex12 + 3 * i);
  EpaOut<float> e;
  if (core) e = epa_penetration<float>(s, [&](V3<float> d) { return pr.core_support(d).w; });
  else e = epa_penetration<float>(s, [&](V3<float> d) { return pr.full_support(d); });
  out4[0] = e.normal.x; out4[1] = e.normal.y; out4[2] = e.normal.z; out4[3] = e.depth;
  *n_vertices = e.vertices;
  *iterations = e.iterations;
  return e.status;
}

}  // extern "C"
