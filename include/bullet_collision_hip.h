/* bullet_collision_hip.h -- C ABI of the MI355X (gfx950) collision hot path:
 * broadphase (overlapping AABB pairs) and convex narrowphase (contact
 * manifolds of at most four points) for the bullet-collision crate of
 * CrazyboyQCD/bullet-physics.
 *
 * Plain C types only: float / uint32 / uint8 arrays and sizes.  Every entry
 * point returns a BCH_* status; on failure bch_last_error() describes it.
 * Host pointers are caller-owned and only read/written during the call.
 * One context per host thread (the reference is single-threaded:
 * `&mut self` everywhere); calls on one context are stream-ordered.
 * A context holds two independent states: the resident world (bch_world_*,
 * bch_sd_*, bch_depen_*) and the batch state (the proxy broadphase
 * bch_bp_*, bch_collide_batch*, the mesh and sweep batches).  Calls on one
 * never change the other, so one context can serve a whole PhysicsWorld.
 *
 * Conventions (crates/bullet-collision/src/narrowphase/mod.rs:21-23):
 *   - a pair (i, j) always has i < j; pair lists are sorted lexicographically
 *     (broadphase/mod.rs:103-118, identical to BruteForceBroadPhase);
 *   - for pair (i, j) body i is shape A and body j is shape B: the contact
 *     normal points from A to B, `position` is the witness point on B,
 *     penetration > 0 is overlap and < 0 a separation inside the contact band;
 *   - a contact is 7 floats: position xyz, normal xyz, penetration.
 *
 * Shapes (crates/bullet-collision/src/shapes):
 *   BCH_SPHERE  params = radius
 *   BCH_BOX     params = half extents x, y, z
 *   BCH_CAPSULE params = radius, half height (segment along local Y)
 *   BCH_HULL    vertices hull_vertices[hull_first[i] .. + hull_count[i]]
 *   BCH_PLANE   params = normal x, y, z (normalised on upload), constant
 * Orientation quaternions are (x, y, z, w).
 */
#ifndef BULLET_COLLISION_HIP_H
#define BULLET_COLLISION_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BCH_OK 0
#define BCH_OVERFLOW 1       /* output capacity too small: *count holds the size needed */
#define BCH_INVALID_ARG 2
#define BCH_HIP_ERROR 3
#define BCH_NO_DEVICE 4

#define BCH_SPHERE 0
#define BCH_BOX 1
#define BCH_CAPSULE 2
#define BCH_HULL 3
#define BCH_PLANE 4

typedef struct bch_context bch_context;

/* Bodies in structure-of-arrays form (host or device pointers, as the entry
 * point states).  hull_* may be NULL when no body is a hull. */
typedef struct bch_bodies {
  uint32_t n;
  const float* position;      /* [3n] */
  const float* orientation;   /* [4n] x, y, z, w */
  const uint8_t* shape_type;  /* [n] BCH_SPHERE .. BCH_PLANE */
  const float* shape_params;  /* [4n] */
  const uint32_t* hull_first; /* [n] first vertex of a hull body */
  const uint32_t* hull_count; /* [n] vertex count of a hull body */
  const float* hull_vertices; /* [3 * n_hull_vertices] local-space xyz */
  uint32_t n_hull_vertices;
} bch_bodies;

/* ------------------------------------------------------------- context -- */
/* No reference counterpart: the reference has no device.  Creates a context
 * on HIP device `device` with its own stream. */
int bch_create(int device, bch_context** out);
void bch_destroy(bch_context* ctx);
const char* bch_last_error(const bch_context* ctx);
int bch_synchronize(bch_context* ctx);
/* Device times (ms) of the last bch_world_collide pass:
 * [0] world AABBs, [1] tree build (Morton codes, sort, nodes, refit),
 * [2] tree traversal (pair search), [3] pair list (planes, compaction,
 * sort, shape-class split), [4] the fork of the narrowphase onto a second
 * stream (near zero), [5] closed-form and GJK pairs, with and without a
 * hull (k_np on the second stream beside k_np_hull), [6] EPA (all its
 * kernels), [7] contact compaction; then [8] the four-pair EPA kernel
 * k_epa_quad alone (part of [6]) and [9] the hull GJK kernel k_np_hull alone
 * (part of [5], beside the second stream).  out_ms holds 10 floats. */
int bch_last_timings(bch_context* ctx, float* out_ms);

/* No reference counterpart: how the EPA entries of the last world pass on
 * this context were routed, read from the device (waits for the context
 * stream): out3[0] entries queued (pairs whose cores overlap) for the
 * four-pair kernel, [1] entries it handed to the one-pair kernel, [2]
 * entries whose polytope outgrew the one-pair kernel's store and reran with
 * the 2048-face store.  Results never depend on the route. */
int bch_last_epa_routes(bch_context* ctx, uint32_t* out3);

/* No reference counterpart: launch-shape knobs of the device passes.  They
 * change how work is spread over the GPU, never the results.
 *   BCH_TUNE_HULL_LANES  lanes per GJK pair with a hull: 4, 8, 16 (16)
 *   BCH_TUNE_TRAVERSE_BLOCK  threads per tree-walk block: 64, 128, 256, 512 (256)
 *   BCH_TUNE_EPA_GRID    EPA blocks of one wavefront (32768)
 *   BCH_TUNE_NP_GRID     narrowphase blocks per shape class (2048)
 *   BCH_TUNE_EPA_WAVE_FACES  faces the one-pair EPA keeps before it hands an
 *                        entry to the 2048-face overflow kernel: 16..256
 *                        (256).  Results do not depend on it; lowering it
 *                        only exercises the overflow route (tests).
 *   BCH_TUNE_SORT_TILE   items per radix-sort tile (histogram and scatter
 *                        block of every sort): 2048 or 4096 (2048) */
enum {
  BCH_TUNE_HULL_LANES = 1,
  BCH_TUNE_TRAVERSE_BLOCK = 2,
  BCH_TUNE_EPA_GRID = 3,
  BCH_TUNE_NP_GRID = 4,
  BCH_TUNE_EPA_WAVE_FACES = 5,
  BCH_TUNE_SORT_TILE = 6
};
int bch_set_tuning(bch_context* ctx, int knob, int value);
int bch_get_tuning(bch_context* ctx, int knob, int* value);
/* Pick the knobs for the bodies loaded by bch_world_set_bodies: times the
 * pass (median of `trials`, device time) over hull lanes, EPA grid,
 * narrowphase grid and tree-walk block, one knob at a time, and keeps
 * the fastest; *out_ms (may be NULL) receives the best pass time. */
int bch_autotune(bch_context* ctx, float aabb_margin, int trials, float* out_ms);

/* ------------------------------------------ BroadPhase trait (proxies) -- */
/* broadphase/mod.rs:94-130.  Slots follow the reference: create_proxy reuses
 * the lowest free slot (brute_force.rs:26-34); set_aabb / destroy on a dead
 * slot are no-ops (axis_sweep3.rs:424-426). */
int bch_bp_create_proxy(bch_context* ctx, const float aabb_min[3], const float aabb_max[3], uint32_t* out_proxy);
int bch_bp_set_aabb(bch_context* ctx, uint32_t proxy, const float aabb_min[3], const float aabb_max[3]);
int bch_bp_destroy_proxy(bch_context* ctx, uint32_t proxy);
/* *live = 0 for a dead or unknown slot (BroadPhase::proxy_aabb -> None). */
int bch_bp_proxy_aabb(bch_context* ctx, uint32_t proxy, float aabb_min[3], float aabb_max[3], int* live);
/* Batched set_aabb for `count` proxies (the world's per-step AABB refresh,
 * world/step.rs:349-363). */
int bch_bp_set_aabbs(bch_context* ctx, uint32_t count, const uint32_t* proxies, const float* mins,
                     const float* maxs);
/* compute_pairs_into (broadphase/mod.rs:103-118): every live (i, j), i < j,
 * whose stored AABBs overlap (closed intervals), sorted; out_pairs[2k] = i,
 * out_pairs[2k+1] = j.  BCH_OVERFLOW when capacity < count. */
int bch_bp_compute_pairs(bch_context* ctx, uint32_t* out_pairs, uint64_t capacity, uint64_t* out_count);
/* aabb_candidates (broadphase/mod.rs:126-129): live proxies overlapping the
 * query box, ascending. */
int bch_bp_aabb_candidates(bch_context* ctx, const float aabb_min[3], const float aabb_max[3], uint32_t* out,
                           uint64_t capacity, uint64_t* out_count);

/* ------------------------------------------------- narrowphase (batch) -- */
/* raycast_candidates (broadphase/mod.rs:119-124, brute_force.rs:70-72):
 * live proxies whose AABB the segment ray_from -> ray_to crosses (slab test
 * of broadphase/mod.rs:153-183), ascending ids; BCH_OVERFLOW as above. */
int bch_bp_raycast_candidates(bch_context* ctx, const float ray_from[3], const float ray_to[3], uint32_t* out,
                              uint64_t capacity, uint64_t* out_count);
/* collide_seeded (narrowphase/mod.rs:63-286) for every pair: shape A =
 * pairs[2k], shape B = pairs[2k+1], seeds[3k..] an optional GJK warm-start
 * direction (NULL = None for every pair; a NaN x component = None for that
 * pair).  out_count[k] = number of contact points (0 =
 * None); out_points[28k + 7c ..] = contact c.  Host pointers. */
int bch_collide_batch(bch_context* ctx, const bch_bodies* bodies, const uint32_t* pairs, uint64_t n_pairs,
                      const float* seeds, uint8_t* out_count, float* out_points);
/* bch_collide_batch on the shapes of the last bch_collide_batch call with
 * new poses (position[3n], orientation[4n], n = that call's body count):
 * only the poses are uploaded.  The per-pair loop of step.rs:445-854 sees
 * the same shapes every step, so only this call runs per step.  Host
 * pointers; BCH_INVALID_ARG before any bch_collide_batch. */
int bch_collide_batch_poses(bch_context* ctx, const float* position, const float* orientation, const uint32_t* pairs,
                            uint64_t n_pairs, const float* seeds, uint8_t* out_count, float* out_points);

/* ------------------------------------------------- triangle meshes -- */
/* A static triangle mesh (TriangleMesh::new, shapes/triangle_mesh.rs:84-123):
 * vertices [3 * n_vertices], triangles [3 * n_triangles] vertex indices.  The
 * BVH is built as the reference builds it (shapes/bvh.rs:170-269), so
 * candidate triangles come in the reference's order.  *out_mesh = its id. */
int bch_mesh_add(bch_context* ctx, const float* vertices, uint32_t n_vertices, const uint32_t* indices,
                 uint32_t n_triangles, uint32_t* out_mesh);
/* Drop every mesh (ids restart at 0). */
int bch_mesh_clear(bch_context* ctx);
/* collide_seeded's convex-vs-mesh arms (narrowphase/mod.rs:177-215,
 * convex_mesh.rs:24-145, without internal-edge info) for a batch: pair k =
 * convex body pairs[2k] of `bodies` against mesh pairs[2k+1] at mesh pose
 * mesh_poses[7k ..] (position xyz, orientation xyzw).  mesh_first[k] != 0
 * (array may be NULL) makes the mesh shape A: the manifold is flipped.
 * Contact band = min(convex, mesh) contact_breaking_threshold.  seeds as in
 * bch_collide_batch.  out_count[k] <= 4 points (ContactManifold::reduce,
 * manifold.rs:76-110), out_points[28k + 7c ..] as in bch_collide_batch.
 * Plane bodies are rejected (plane-vs-mesh is not part of this path).  Host
 * pointers. */
int bch_collide_mesh_batch(bch_context* ctx, const bch_bodies* bodies, const uint32_t* pairs, const float* mesh_poses,
                           const uint8_t* mesh_first, uint64_t n_pairs, const float* seeds, uint8_t* out_count,
                           float* out_points);

/* ------------------------------------------------------ CCD sweeps -- */
/* Sweeps (narrowphase/convex_sweep.rs) for a batch of queries, one per call
 * of the reference's sweep functions: query k moves from from[3k ..] to
 * to[3k ..] against body target[k] of `bodies` at its pose.  caster == NULL:
 * swept_sphere_cast (convex_sweep.rs:37-57) of a sphere of radius[k], as
 * the CCD pass of DiscreteDynamicsWorld::step calls it
 * (bullet-dynamics/src/world/step.rs:1141-1148).  Else convex_cast
 * (convex_sweep.rs:370-385) of body caster[k]'s shape at its orientation, as
 * convex_sweep_closest_filtered calls it (bullet-dynamics/src/world/
 * query.rs:210-218); cast shapes must be convex.  Targets may be any convex
 * shape or a plane.  out_hit[k] = 1 for Some(hit); out_sweep[7k ..] =
 * fraction, normal (from the target toward the caster), point on the
 * target (zeros for None).  Host pointers. */
int bch_sweep_batch(bch_context* ctx, const bch_bodies* bodies, const uint32_t* target, const uint32_t* caster,
                    const float* from, const float* to, const float* radius, uint64_t n, uint8_t* out_hit,
                    float* out_sweep);
/* The mesh arms of the same functions (sweep_mesh, convex_sweep.rs:294-344;
 * sweep_convex_mesh, :480-524): query k against mesh mesh_of[k] at
 * mesh_poses[7k ..] (position xyz, orientation xyzw); the earliest triangle
 * hit (the first in BVH order on equal fractions).  bodies may be NULL when
 * caster is NULL.  Outputs as in bch_sweep_batch. */
int bch_sweep_mesh_batch(bch_context* ctx, const bch_bodies* bodies, const uint32_t* mesh_of, const float* mesh_poses,
                         const uint32_t* caster, const float* from, const float* to, const float* radius, uint64_t n,
                         uint8_t* out_hit, float* out_sweep);

/* ---------------------------------------------- world-resident pipeline -- */
/* The per-step detection pass of PhysicsWorld (world/step.rs:349-416 +
 * 445-854 for rigid pairs): world AABBs padded by `aabb_margin`, the sorted
 * overlapping pair set, collide_seeded(body i, body j, None) per pair.
 * Bodies stay resident in device memory between passes. */

/* Upload (copy) all bodies from host memory; a batch of independent worlds
 * passes world_offsets[n_worlds + 1] (bodies of world w are
 * world_offsets[w] .. world_offsets[w+1]; pairs never cross worlds), or NULL
 * and n_worlds = 1 for one world. */
int bch_world_set_bodies(bch_context* ctx, const bch_bodies* bodies, uint32_t n_worlds,
                         const uint32_t* world_offsets);
/* Per-step pose refresh from host memory (positions [3n], orientations [4n]). */
int bch_world_set_poses(bch_context* ctx, const float* position, const float* orientation);
/* Per-step pose refresh from device memory on the same device. */
int bch_world_set_poses_device(bch_context* ctx, const float* d_position, const float* d_orientation);
/* Pose refresh at the step's predicted positions (step.rs:340-343,
 * body.rs:1175-1181): position + linear_velocity [3n] * dt for bodies that
 * are active (BCH_BODY_ACTIVE of bch_world_set_body_flags; every body when
 * no flags are set), the position as given for the others; orientations as
 * given.  Static bodies carry zero velocity.  Host / device arrays. */
int bch_world_set_poses_predicted(bch_context* ctx, const float* position, const float* orientation,
                                  const float* linear_velocity, float dt);
int bch_world_set_poses_predicted_device(bch_context* ctx, const float* d_position, const float* d_orientation,
                                         const float* d_linear_velocity, float dt);
/* Run one pass on the context stream; returns after enqueueing plus one
 * small device->host read of the pair count.  *out_pairs / *out_contacts
 * receive the counts. */
int bch_world_collide(bch_context* ctx, float aabb_margin, uint64_t* out_pairs, uint64_t* out_contacts);
/* Copy the last pass's results to host memory: pairs [2P], per-pair contact
 * offsets [P+1] (contacts of pair k are offsets[k] .. offsets[k+1]) and
 * contacts [7C].  Any pointer may be NULL to skip that array. */
int bch_world_results(bch_context* ctx, uint32_t* pairs, uint32_t* contact_offsets, float* contacts);
/* No reference counterpart (diagnostic): the route every pair of the last
 * pass took through collide_seeded's dispatch (narrowphase/mod.rs:99-285),
 * one byte per pair in pair order: 0 none (plane-plane), 1 sphere-sphere,
 * 2 sphere-box, 3 plane-convex, 4 box-box, 5 convex-convex without a hull,
 * 6 convex-convex with a hull; +8 when the pair went on to EPA
 * (convex_convex.rs:117-154). */
int bch_world_routes(bch_context* ctx, uint8_t* routes);
/* ------------------------------------------------- signed distance -- */
/* Signed distance of each pair of the bodies loaded by bch_world_set_bodies
 * (at their current poses): the margin-core walk of gjk_contact
 * (narrowphase/gjk.rs:171-288) with threshold = +inf, so a separated pair of
 * any distance ends in its Band arm, and EPA plus the contact of
 * convex_convex.rs:117-154 for overlapping cores.  A walk stopped by its
 * iteration cap reports its closest iterate like the Band arm, and so does a
 * walk that proved a core gap above 1e-4 m (one iterate's support plane) and
 * then ended on a simplex enclosing the origin, which only rounding can
 * produce (the reference's f32 walk would go on to EPA with it and report a
 * false, often deep, penetration), and an EPA whose core depth exceeds the
 * walk's own bound (the least support-plane distance -delta / |v| of its
 * iterates) by more than 1e-4 m, which only a polytope corrupted by rounding
 * can give.  All pointers
 * are DEVICE pointers; the calls are ordered on the context stream and do
 * not synchronise.
 *   d_pairs  [2P] body indices (i, j): shape A = body i, shape B = body j
 *   d_out    [10P] per pair: sd (> 0 separated, < 0 penetrating), unit normal
 *            A -> B, witness on A, witness on B (pa = pb - n sd)
 *   d_status [P] or NULL: 0 walk, 1 EPA, 2 capped walk, 3 touching cores (no
 *            normal), 4 EPA unconverged, 5 walk ended by rounding after it
 *            proved a gap (closest iterate), 6 EPA contradicted by the
 *            walk's depth bound (closest iterate); +8 EPA depth from an
 *            edge, +16 a
 *            face vertex not split (backward only); 255 pair index out of
 *            range (outputs NaN).
 * No reference counterpart beyond the walks it reuses. */
int bch_sd_forward(bch_context* ctx, const uint32_t* d_pairs, uint64_t n_pairs, float* d_out, uint8_t* d_status);
/* Adjoint of bch_sd_forward at the same poses: d_grad_out [10P] = dL/d(out),
 * d_grad_pose [14P] = dL/d(position A xyz, quaternion A xyzw, position B,
 * quaternion B) per pair (the caller sums over pairs sharing a body).  The
 * walks are recomputed, nothing is stored between the two calls. */
int bch_sd_backward(bch_context* ctx, const uint32_t* d_pairs, uint64_t n_pairs, const float* d_grad_out,
                    float* d_grad_pose, uint8_t* d_status);
/* Order this context's work on `stream` (a hipStream_t of the same device,
 * e.g. the caller's framework stream; NULL is the device's null stream).
 * The previous stream is drained first. */
int bch_set_stream(bch_context* ctx, void* stream);
/* Return to the context's own stream (the previous one is drained first). */
int bch_reset_stream(bch_context* ctx);

/* Device pointers of the last pass's results (valid until the next pass). */
int bch_world_results_device(bch_context* ctx, const uint32_t** d_pairs, const uint32_t** d_contact_offsets,
                             const float** d_contacts);

/* ------------------------------------------------ captured world pass -- */
/* No reference counterpart: the world pass of bch_world_collide for callers
 * that step many frames (the per-step pair and contact pass of
 * world/step.rs:445-854 over every world of a batch) without a host round
 * trip.  The pass is run in capacity mode, with buffers for pair_capacity
 * pairs and contact_capacity contacts, counts and overflow flags kept on the
 * device, and captured in a HIP graph.  Before capturing, one uncaptured
 * pass at the current poses sizes the buffers;
 * its counts are returned like bch_world_replay_counts (BCH_OVERFLOW if that
 * pass already overflowed: capture again with larger capacities).  Results
 * are identical to bch_world_collide's. */
int bch_world_capture(bch_context* ctx, float aabb_margin, uint64_t pair_capacity, uint64_t contact_capacity);
/* Launch the captured pass on the context stream (one hipGraphLaunch, no
 * synchronisation), reading the poses the bodies hold when it runs.
 * BCH_INVALID_ARG if nothing is captured or a pass buffer was reallocated
 * since (e.g. an eager pass with more pairs than the capacity): capture
 * again. */
int bch_world_replay(bch_context* ctx);
/* Wait for the last replayed (or capture) pass and read its counts:
 * *n_pairs / *n_contacts are what the buffers hold, the sizes
 * bch_world_results then writes; *found_pairs / *found_contacts what the
 * pass found (any pointer may be NULL).  The two agree unless the result is
 * BCH_OVERFLOW: a capacity was exceeded, the
 * held results are incomplete, capture again with capacities above the
 * found counts. */
int bch_world_replay_counts(bch_context* ctx, uint64_t* n_pairs, uint64_t* n_contacts, uint64_t* found_pairs,
                            uint64_t* found_contacts);
/* Free the captured pass (bch_world_set_bodies and bch_destroy do too). */
int bch_world_release_graph(bch_context* ctx);

/* ------------------------------------------- persistent contact manifolds -- */
/* The rigid-pair arm of the world step's pair pass with persistent
 * manifolds (PersistentManifold, crates/bullet-collision/src/manifold.rs:
 * 124-386, driven as in crates/bullet-dynamics/src/world/step.rs:660-854),
 * for the bodies loaded by bch_world_set_bodies at their current poses
 * (positions are the step's predicted positions).  Per broadphase pair:
 *   - group / mask filter (step.rs:660-664): the pair is dropped;
 *   - both bodies inactive, or both static / kinematic: the cached manifold
 *     is kept as it is, nothing is emitted (step.rs:683-713);
 *   - a cached manifold within the redetect drift (5 mm, 0.01 rad) is
 *     refreshed and its points emitted, unless a body is fast and a point
 *     is deeper than linear_slop (step.rs:715-776);
 *   - otherwise collide_seeded runs with the cache's seed; its manifold is
 *     captured (merge, impulse inheritance, deepest four) and the refreshed
 *     points emitted, or the entry is dropped when it finds nothing
 *     (step.rs:777-852).
 * Emitted impulses are the cached ones times warm_start (step.rs:751, 821).
 * The cache lives on the device between calls; entries of pairs that are
 * no longer broadphase pairs are dropped (the prune of step.rs:1006-1010).
 * Per-body flags: BCH_BODY_STATIC (static or kinematic), BCH_BODY_ACTIVE,
 * BCH_BODY_FAST (settled speed^2 above the redetect speed, world.rs:62). */
#define BCH_BODY_STATIC 1
#define BCH_BODY_ACTIVE 2
#define BCH_BODY_FAST 4
/* Routes per pair. */
#define BCH_ROUTE_FILTERED 0
#define BCH_ROUTE_SLEEPING 1
#define BCH_ROUTE_STATIC 2
#define BCH_ROUTE_SERVED 3   /* refreshed from the cache */
#define BCH_ROUTE_DETECTED 4 /* collide_seeded + capture */
#define BCH_ROUTE_REMOVED 5  /* collide_seeded found nothing: entry dropped */
/* Per-body flags [n] and collision group | mask << 16 [n] (host arrays; NULL
 * = every body active and dynamic and slow / everything collides).  Kept
 * until the next call or bch_world_set_bodies.  While flags are set, a body
 * without BCH_BODY_ACTIVE keeps the broadphase AABB of the previous pass
 * (step.rs:348-353: sleeping bodies keep theirs) in every pass. */
int bch_world_set_body_flags(bch_context* ctx, const uint8_t* flags, const uint32_t* group_mask);
/* One pair pass at the current poses (synchronous).  *out_pairs = broadphase
 * pairs, *out_points = emitted points. */
int bch_world_manifold_pass(bch_context* ctx, float aabb_margin, float linear_slop, float warm_start,
                            uint64_t* out_pairs, uint64_t* out_points);
/* The last pass's results (any pointer may be NULL): pairs [2P], route [P],
 * per-pair point offsets [P+1] and the points [12 * n_points]: position on
 * B xyz, normal A->B xyz, penetration, normal impulse (warm-started),
 * lifetime (uint32 bits), A-side anchor xyz (position A + orientation A *
 * the cached local anchor, for friction anchoring). */
int bch_world_manifold_results(bch_context* ctx, uint32_t* pairs, uint8_t* route, uint32_t* point_offsets,
                               float* points);
/* Post-solve writeback (step.rs:978-1005): impulses [n_points] of the last
 * pass's points, in results order, into their cached points. */
int bch_world_manifold_impulses(bch_context* ctx, const float* impulses);
/* No reference counterpart (diagnostic): the cache after the last pass, per
 * pair: valid [P] and entries [64 P] floats (4 points x (local anchor A xyz,
 * penetration, local anchor B xyz, impulse, normal xyz, lifetime bits),
 * detect position A, threshold, detect orientation A, detect position B,
 * count | initialised << 8 bits, detect orientation B). */
int bch_world_manifold_cache(bch_context* ctx, uint8_t* valid, float* entries);
/* Drop every cached manifold. */
int bch_world_manifold_reset(bch_context* ctx);

/* ------------------------------------------ captured de-penetration step -- */
/* No reference counterpart: gradient descent on the penetration penalty
 * L = sum of max(0, -sd)^2 over the pairs the broadphase finds (AABBs padded
 * by aabb_margin), sd the signed distance of bch_sd_forward, for the bodies
 * loaded by bch_world_set_bodies.  One step, on the device with no host
 * round trip:
 *   broadphase -> signed distance of every pair -> record (loss, largest
 *   penetration, pair and penetrating-pair counts) -> adjoint of the
 *   penetrating pairs (bch_sd_backward with dL/dsd = -2 max(0, -sd)) ->
 *   per-body gradient sum -> d_position -= lr dL/dposition,
 *   d_orientation = normalize(d_orientation - lr dL/dorientation).
 * d_position [3n] / d_orientation [4n] are DEVICE arrays, read at the start
 * of each step and updated in place; they must outlive the capture.  The
 * step is captured in a HIP graph after one uncaptured step without the
 * update (poses unchanged) has sized the buffers for pair_capacity pairs;
 * *found_pairs (may be NULL) receives that step's pair count, BCH_OVERFLOW
 * when it exceeds the capacity.  history_capacity steps are recorded. */
int bch_depen_capture(bch_context* ctx, float aabb_margin, float lr, float* d_position, float* d_orientation,
                      uint64_t pair_capacity, uint32_t history_capacity, uint64_t* found_pairs);
/* Launch the captured step `steps` times on the context stream (no
 * synchronisation).  BCH_INVALID_ARG when nothing is captured or a buffer
 * was reallocated since. */
int bch_depen_replay(bch_context* ctx, uint32_t steps);
/* Wait, then copy the records of the steps run since the capture (or the
 * last reset) into the arrays (any may be NULL; at most `capacity` each):
 * loss, largest penetration, broadphase pairs found, penetrating pairs.
 * *count = steps run.  reset != 0 starts a new history.  BCH_OVERFLOW when a
 * step found more pairs than the capacity (its gradient missed pairs:
 * capture again with a larger capacity) or ran past the history capacity. */
int bch_depen_history(bch_context* ctx, float* loss, float* max_pen, uint32_t* n_pairs, uint32_t* n_pen,
                      uint32_t capacity, uint32_t* count, int reset);
/* Wait, then free the captured step (it holds the pose array pointers;
 * bch_world_set_bodies and bch_destroy free it too). */
int bch_depen_release(bch_context* ctx);

/* --------------------------------------------- no allocation (static) -- */
/* No counterpart in the reference: this is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  bch_collide_batch with device arrays:
 * every pointer inside d_bodies and every array argument is a device pointer
 * (d_bodies itself, with n and n_hull_vertices, is host memory); d_pairs
 * u32[2 n_pairs], d_seeds f32[3 n_pairs] or NULL, d_out_count u8[n_pairs],
 * d_out_points f32[28 n_pairs] in bch_collide_batch's layout.  The working
 * memory is one static device array of BCH_STATIC_BYTES per GPU, held until
 * the process ends; the call makes no hipMalloc, hipFree or allocating
 * library call.  Inputs beyond BCH_STATIC_MAX_PAIRS pairs,
 * BCH_STATIC_MAX_BODIES bodies or BCH_STATIC_MAX_HULL_VERTICES hull vertices,
 * and bad shape types, hull ranges or pair indices (checked on the device
 * first), return BCH_INVALID_ARG.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other threads, other contexts, other
 * streams) race.  The outputs are written in stream order. */
#define BCH_STATIC_MAX_PAIRS 32768u
#define BCH_STATIC_MAX_BODIES 65536u
#define BCH_STATIC_MAX_HULL_VERTICES 131072u
#define BCH_STATIC_BYTES 16074752u
int bch_collide_batch_static(bch_context* ctx, const bch_bodies* d_bodies, const uint32_t* d_pairs, uint64_t n_pairs,
                             const float* d_seeds, uint8_t* d_out_count, float* d_out_points);

#ifdef __cplusplus
}
#endif
#endif
