/*
 * ptg.h -- C ABI of libptg_hip.so, the gfx950 tree-gravity library.
 *
 * The reference (pytreegrav) has no C ABI: its operator interface is the Python API in
 * src/pytreegrav/frontend.py and the Octree jitclass in src/pytreegrav/octree.py.  Each entry point
 * below replaces the compute behind one of those Python functions; the Python mirror in
 * pytreegrav_amd/frontend.py binds them with ctypes.
 *
 * Conventions
 *   - Every function returns a ptg_status: 0 on success, a negative code on failure; the message
 *     is available from ptg_last_error() (thread-local).  The library never aborts or prints.
 *   - All array arguments are device pointers owned by the caller (the Python side allocates them
 *     as torch tensors).  Scratch memory is a caller-provided workspace whose size is returned by
 *     the matching *_workspace_bytes function.
 *   - Work is enqueued asynchronously on the given HIP stream (NULL = default stream).  Host
 *     synchronisations: ptg_build_sort (finiteness and forest checks, collision count, node count)
 *     and ptg_walk_check (reads the overflow flag).  Every other call only enqueues work.
 *   - Element numbering follows the reference Octree (octree.py:8-28): particles occupy
 *     [0, n) in Morton order, the root is element n, internal nodes follow.  Links are int32.
 */
#ifndef PTG_H
#define PTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int ptg_status;
#define PTG_OK 0
#define PTG_ERR_ARG (-1)      /* invalid argument */
#define PTG_ERR_HIP (-2)      /* HIP runtime error */
#define PTG_ERR_OVERFLOW (-3) /* a walk traversal stack exceeded its spill capacity */
#define PTG_ERR_WORKSPACE (-4)/* workspace smaller than *_workspace_bytes asked for */

/* field bits, packed in this order in every output row: potential, acceleration(3), tidal(9)
 * (grouped_treewalk.py:48-52, _field_width :264-266) */
#define PTG_POT 1
#define PTG_ACC 2
#define PTG_TIDAL 4

#define PTG_F32 0
#define PTG_F64 1

int ptg_version(void);
const char *ptg_last_error(void);

/* ------------------------------------------------------------------------------------------
 * Tree build: replaces Octree.__init__ radix path (octree.py:668-723), i.e. BuildTreeRadix
 * (:817-1091) with _bounding_cube (:539-574), _morton_keys (:69-98), _radix_argsort (:101-132),
 * _permute_particles (:577-603), RekeySlice (:1093-1126), BuildBucket (:1128-1175), and the
 * moment passes RollupMoments (:199-231) / RollupQuadrupoles (:252-292).
 * Two calls, because the node count is data dependent:
 *   ptg_build_sort   bbox, keys, sort, collision re-keying, node counting (one host sync)
 *   ptg_build_tree   node emission, links, moments, into caller arrays sized from info.nnodes
 * ------------------------------------------------------------------------------------------ */
typedef struct {
    int64_t n;              /* particles */
    int64_t nnodes;         /* internal nodes including the root */
    int32_t has_collisions; /* some particles shared a 63-bit key and were re-keyed */
    int32_t has_coincident; /* some positions were bit-identical (bucket chains), HasCoincidentPoints */
    int32_t max_level;      /* deepest node level (root = 0) */
    int32_t nonfinite;      /* set when the input held a NaN or infinity (the call then fails) */
    double center[3];       /* root cube centre */
    double size;            /* root cube side */
} ptg_build_info;

typedef struct {
    double *coordinates;      /* (n+nnodes, 3): particle positions, node COM (geometric centre if mass 0) */
    double *masses;           /* (n+nnodes) */
    double *softenings;       /* (n+nnodes): particle h, node max h */
    double *sizes;            /* (n+nnodes): node side; particle rows are not written (zero) */
    double *deltas;           /* (n+nnodes): |COM - geometric centre|; particle rows not written */
    double *quadrupoles;      /* (n+nnodes, 9) traceless quadrupoles of the nodes, or NULL for a monopole
                               * tree; particle rows are not written (a particle's quadrupole is zero) */
    int32_t *first_subnode;   /* (n+nnodes), -1 for particles */
    int32_t *next_branch;     /* (n+nnodes), -1 terminates the walk */
    int64_t *treewalk_indices;/* (n): input index of each Morton slot */
    int32_t *nchild;          /* (n+nnodes): number of children (0 for particles) */
    int32_t *child_base;      /* (n+nnodes): walk slot of the first child (sibling-contiguous layout) */
    int32_t *slot;            /* (n+nnodes): walk slot of each element (the root is slot 0) */
    uint64_t *keys;           /* (n) or NULL: 63-bit Morton key of each particle slot (of its system's cube) */
    double *node_stats;       /* (nnodes, 5) or NULL (build with quadrupoles and moments): per node |Q|_F, the
                               * largest child mass, sum m d^2, sum m d^3 and the largest d over its
                               * children (d = child centre to the node's centre of mass); the learned
                               * opening radii read these instead of walking each node's links */
    int32_t *node_count;      /* device int32 or NULL.  Set for a tree from ptg_build_deferred: the node
                               * arrays are sized for a capacity, the build writes the number of nodes
                               * here (0 if the tree did not fit), and every call that takes nnodes with
                               * this tree treats nnodes as the capacity and reads the count here */
} ptg_tree_arrays;

/* Forest: several independent systems (no reference counterpart; the reference builds one Octree
 * per call).  Each system gets its own bounding cube, Morton keys and tree, exactly as if it were
 * built alone, but all systems are built, packed, grouped and walked by the same launches.
 * System s owns input particles [offsets[s], offsets[s+1]), which after the build are also its
 * Morton slots (the tree stays system-major), and its root is element roots[s], at walk slot s.
 * Every system needs at least two particles.  A NULL forest means one tree; a forest with
 * nsys == 1 is the same tree and also fills cubes/roots. */
typedef struct {
    int64_t nsys;           /* number of systems */
    const int64_t *offsets; /* device (nsys + 1): 0 = offsets[0] < ... < offsets[nsys] = n */
    int32_t *roots;         /* device (nsys): element index of each root; written by ptg_build_tree */
    double *cubes;          /* device (nsys, 4): root cube centre x, y, z and side; written by ptg_build_sort */
} ptg_forest;

/* ------------------------------------------------------------------------------------------
 * Walk tree: the sibling-contiguous packed copy of the tree the walk reads (one slot per
 * element; the children of a node occupy nchild consecutive slots from child_base).  Coordinates
 * are stored relative to `origin` (node-relative precision for FP32 packs; pass 0 for FP64).  For
 * a forest, `origin` is ignored: FP32 packs are relative to each system's cube centre and FP64
 * packs absolute.
 * Per slot: xyzm (4 reals: x, y, z, mass), aux (4 reals: softening, size, delta, 0),
 * kids (2 int32: first child slot, child count; count 0 marks a particle),
 * quad (8 reals: xx, xy, xz, yy, yz, zz, 0, 0) when quadrupole != 0.
 * ------------------------------------------------------------------------------------------ */
typedef struct {
    void *xyzm;
    void *aux;
    int32_t *kids;
    void *quad; /* NULL for a monopole pack */
    int64_t nslots;
    int32_t prec; /* PTG_F32 / PTG_F64 */
    int32_t has_quad;
    double origin[3];
} ptg_walk_tree;

/* Phase timing (no reference counterpart; diagnostics for the build's per-phase HBM rates): the
 * PTG_PHASE_COUNT hipEvent_t handles registered here are recorded on the build stream at these
 * points of every later ptg_build_sort / ptg_build_tree call of the calling thread, until
 * ptg_build_phase_events(NULL, 0).  Boundaries: START (build_sort entered), BBOX (bounding cube),
 * KEYS (Morton keys and particle records), SORT (radix sort), COUNT_NODES (shared levels, collision
 * re-keying, node count), TREE_START (build_tree entered), PERMUTE (Morton-order particle gather),
 * NODES (node order, geometry, FirstSubnode/NextBranch), LINKS (child counts, parents, walk slots),
 * MOMENTS (mass, centre of mass, quadrupoles), PACK (FP32 walk copy).  (The bounding cube and key
 * marks are recorded by single-tree builds.) */
#define PTG_PHASE_START 0
#define PTG_PHASE_BBOX 1
#define PTG_PHASE_KEYS 2
#define PTG_PHASE_SORT 3
#define PTG_PHASE_COUNT_NODES 4
#define PTG_PHASE_TREE_START 5
#define PTG_PHASE_PERMUTE 6
#define PTG_PHASE_NODES 7
#define PTG_PHASE_LINKS 8
#define PTG_PHASE_MOMENTS 9
#define PTG_PHASE_PACK 10
#define PTG_PHASE_COUNT 11
ptg_status ptg_build_phase_events(void *const *events, int nevents);

size_t ptg_build_workspace_bytes(int64_t n, int64_t nsys);
/* A non-finite position, mass or softening fails with PTG_ERR_ARG and info->nonfinite = 1
 * (ConstructTree's check, frontend.py:123-125).  Pass ptg_build_tree the same pos, m and h and the
 * same workspace: the sort also packs the inputs into the workspace, and the tree's particle slots
 * are gathered from that copy.
 * For a forest, info->center/size describe system 0 (all cubes are in forest->cubes). */
ptg_status ptg_build_sort(const double *pos, const double *m, const double *h, int64_t n, const ptg_forest *forest,
                          void *ws, size_t ws_bytes, ptg_build_info *info, void *stream);
size_t ptg_node_workspace_bytes(int64_t n, int64_t nnodes, int max_level);
/* particles_ready (nullable hipEvent_t): recorded on the stream as soon as the particle slots
 * (coordinates, masses, softenings, treewalk_indices, keys) are written, so a caller can start
 * work that needs only those (target groups) on another stream while nodes and moments are built.
 * pack (nullable): a walk copy to fill at the end of the build, the same as ptg_pack_walk's with the
 * quadrupole norms in aux[3], but written in slot order through the build's slot map (coalesced
 * stores); its ptg_walk_tree is as for ptg_pack_walk (nslots = n + nnodes). */
ptg_status ptg_build_tree(const double *pos, const double *m, const double *h, const ptg_build_info *info,
                          const ptg_forest *forest, int quadrupole, int compute_moments, void *ws, size_t ws_bytes,
                          void *node_ws, size_t node_ws_bytes, const ptg_tree_arrays *tree,
                          const ptg_walk_tree *pack, void *particles_ready, void *stream);

/* ------------------------------------------------------------------------------------------
 * Deferred build (no reference counterpart): ptg_build_sort + ptg_build_tree in one call that
 * never synchronises with the host, so a caller can capture it (with the groups and the walk) in a
 * HIP graph and replay it on new positions of the same particle count.  The node arrays (tree,
 * pack) are sized for n + nnodes_cap elements and the moments run for levels 0..level_cap; the
 * build writes the actual node count to *tree->node_count and everything else to *counts (device).
 * If the tree needs more than nnodes_cap nodes or level_cap levels, or more than
 * PTG_DEFERRED_COLLISION_CAP particles share 63-bit keys, *tree->node_count is 0, counts->overflow
 * says why and the node arrays are left unset (rebuild with larger capacities, or with
 * ptg_build_sort + ptg_build_tree).  Forest offsets are not validated here (ptg_build_sort does).
 * Workspaces: ptg_build_workspace_bytes(n, nsys) and ptg_node_workspace_bytes(n, nnodes_cap,
 * level_cap).  Results equal those of the two-call build bit for bit.
 * ------------------------------------------------------------------------------------------ */
#define PTG_DEFERRED_COLLISION_CAP 8192
#define PTG_OVERFLOW_NODES 1
#define PTG_OVERFLOW_LEVELS 2
#define PTG_OVERFLOW_COLLISIONS 4
#define PTG_OVERFLOW_SORT 8
typedef struct {
    int32_t nnodes;         /* internal nodes including the roots (as computed, even when over capacity) */
    int32_t max_level;      /* deepest node level */
    int32_t has_collisions; /* some particles shared a 63-bit key */
    int32_t has_coincident; /* bucket chains of bit-identical positions */
    int32_t nonfinite;      /* a position, mass or softening was NaN or infinite (the tree is garbage) */
    int32_t overflow;       /* PTG_OVERFLOW_* bits; 0 when the tree is complete */
    int32_t pad[2];
    double center[3];       /* root cube centre (system 0 of a forest) */
    double size;            /* root cube side */
} ptg_build_counts;

ptg_status ptg_build_deferred(const double *pos, const double *m, const double *h, int64_t n,
                              const ptg_forest *forest, int quadrupole, int compute_moments, int64_t nnodes_cap,
                              int level_cap, void *ws, size_t ws_bytes, void *node_ws, size_t node_ws_bytes,
                              const ptg_tree_arrays *tree, const ptg_walk_tree *pack, ptg_build_counts *counts,
                              void *particles_ready, void *stream);

/* Morton permutation of arbitrary points: replaces grouped_treewalk._morton_order (:269-275).
 * Uses a ptg_build_workspace_bytes(n, 1) workspace; asynchronous. */
ptg_status ptg_morton_order(const double *pos, int64_t n, int64_t *order, void *ws, size_t ws_bytes, void *stream);

/* Target groups for the walk.  PTG_GROUPS_RUNS: runs of group_size consecutive sorted targets,
 * the reference's grouping (grouped_treewalk.py:179-182).  PTG_GROUPS_CELLS (no reference
 * counterpart, the FP32 default): groups never cross the boundary of a cell holding more than
 * 4*group_size targets and hold at most group_size targets.  Targets pos[perm[t]] must be in
 * Morton order.  Writes offsets[0..G] (capacity n+1) and the group count G to the device scalar
 * *ngroups; asynchronous.  With a forest the targets are the forest's own particles (perm = its
 * treewalk_indices): runs restart and cells are cut at every system start, and each system's
 * cube defines its cells, so every system is grouped exactly as if it were alone.  keys
 * (nullable): the targets' Morton keys in sorted order when the caller has them (the tree's own
 * particles: ptg_tree_arrays.keys), which skips the bounding cube and key pass. */
#define PTG_GROUPS_RUNS 0
#define PTG_GROUPS_CELLS 1
size_t ptg_groups_workspace_bytes(int64_t n, int64_t nsys);
ptg_status ptg_target_groups(const double *pos, const int64_t *perm, const uint64_t *keys, int64_t n,
                             const ptg_forest *forest, int group_size, int mode, int32_t *offsets, int32_t *ngroups,
                             void *ws, size_t ws_bytes, void *stream);

/* Field-scale sample targets (no reference counterpart: the sample walk that sets the RMS-field
 * unit of the error-bounded and learned criteria).  From the groups (offsets, *ngroups) of
 * ptg_target_groups over a tree's own n particles, takes `samples` groups of every system, evenly
 * spaced over the system's groups (the system's first group + k * its group count / samples), pads
 * each to group_size targets by repeating its last target, and writes their sorted coordinates and
 * softenings (read from the tree's Morton-ordered coordinates / softenings) to row
 * (s * samples + k) * group_size + j of pos (rows x 3) and soft (rows), rows = nsys * samples *
 * group_size.  forest: NULL for a single tree.  Asynchronous; no host synchronisation. */
ptg_status ptg_field_sample(const int32_t *offsets, const int32_t *ngroups, const ptg_forest *forest, int64_t n,
                            int samples, int group_size, const double *coordinates, const double *softenings,
                            double *pos, double *soft, void *stream);


ptg_status ptg_pack_walk(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const ptg_forest *forest,
                         const ptg_walk_tree *wt, void *stream);

/* ------------------------------------------------------------------------------------------
 * Grouped walk: replaces grouped_treewalk.FieldsTarget_grouped (:350-397) with the core
 * (:163-241), the per-interaction kernel (:45-155) and acceptance_criterion (treewalk.py:22-25).
 * Targets are read as pos[perm[t]] for t in Morton order (perm = TreewalkIndices for self
 * gravity, the _morton_order permutation for external targets) and results are written to
 * out[perm[t]], i.e. in the caller's original order: out is (nt, W) float64 with
 * W = 1*POT + 3*ACC + 9*TIDAL, scaled by G.  One wavefront walks each group of group_size
 * consecutive sorted targets (1..256), or, when group_offsets/group_count (device arrays from
 * ptg_target_groups) are given, each group [offsets[g], offsets[g+1]).  exact != 0 evaluates every pair with the reference's
 * branches (r2 > 0 self skip, softened iff r < max(h_s, h_t)); exact == 0 (FP32 fast path) skips
 * those tests where the group geometry proves them false.  counters (nullable, device int64[4])
 * accumulate: [0] element-target interactions, [1] traversal visits, [2] the part of [0] taken
 * through the particle list (near-field pairs whose softening is checked), [3] element-group pairs
 * folded into group expansions (see ptg_criterion).  With a forest (the same
 * descriptor the tree was built and packed with) the targets are the forest's own particles
 * (perm = its treewalk_indices, nt = n, groups that never cross a system, e.g. from
 * ptg_target_groups with the forest or group_size-runs when every system size is a multiple of
 * group_size): each group walks the tree of the system holding its first target.
 * ------------------------------------------------------------------------------------------ */
/* Opening criterion.  PTG_CRIT_THETA: the reference's geometric test (treewalk.py:22-25).
 * PTG_CRIT_QERR (no reference counterpart): a node is accepted only if it also passes the
 * geometric test at `theta` (used as a loose cap) and its dropped quadrupole term, estimated as
 * 3.5 |Q|_F / r_min^4 with r_min the distance to the group's bounding box, is below `eps`
 * (an absolute acceleration tolerance, in units of G=1); needs a tree built with quadrupoles. */
#define PTG_CRIT_THETA 0
#define PTG_CRIT_QERR 1
/* PTG_CRIT_RADIUS (no reference counterpart): a node is accepted only if it also passes the
 * geometric test at `theta` (the guard) and r_min exceeds the opening radius stored in aux[3] of
 * its slot by ptg_open_radii (the learned criterion); eps is unused (it is baked into the radii). */
#define PTG_CRIT_RADIUS 2
/* group_expansion = 1 (0 = off; no reference counterpart; FP32 walk, monopole interactions, potential and/or
 * acceleration): an accepted element of mass M at distance D from the centre of a group's bounding
 * box (R = its half diagonal, x = R / D < 1) whose truncation bound (M / D^2) x^3 (4 - 3x) / (1 - x)^2
 * is below expansion_eps is not evaluated per target but folded into a third-order Taylor expansion
 * of its field about the group centre, evaluated once per target after the traversal.
 * expansion_eps is an absolute acceleration (G = 1), times eps_scale[s] (device, nullable; s = 0 for
 * one tree, the group's system for a forest) when given. */
/* expansion_rel (group expansion; 0 = off): an element is folded only if its truncation bound is
 * also below expansion_rel x (M / D^2), its own field at the group centre, i.e. x = R / D is capped
 * by x^3 (4 - 3x) / (1 - x)^2 < expansion_rel whatever the absolute tolerance allows.
 * variant (no reference counterpart): the kernel variant of the FP32 one-target-per-lane walk
 * (group_size <= 64) for monopole potential/acceleration, 0 = 5 waves per SIMD with interaction lists
 * evaluated at 64 entries, 1 = 4 waves per SIMD, 2 = 6 waves per SIMD with lists at 32 (less LDS per
 * wave); variant 1 gives results bit-identical to 0, variant 2 sums the same terms in another order
 * (FP32 rounding-level differences); the autotuner picks one.  Other walks ignore it. */
typedef struct {
    int32_t criterion;
    int32_t group_expansion;
    double eps;
    double expansion_eps;
    const double *eps_scale;
    int32_t variant;
    float expansion_rel;
} ptg_criterion;

size_t ptg_walk_workspace_bytes(int64_t nt, int group_size);
/* targets_sorted != 0: pos and soft are already in sorted target order (pos[t], soft[t], e.g. the
 * tree's own Morton-ordered coordinates and softenings for self gravity); perm then only places
 * the outputs. */
ptg_status ptg_walk(const ptg_walk_tree *wt, const ptg_forest *forest, const double *pos, const double *soft,
                    const int64_t *perm, int targets_sorted, int64_t nt, int fields, int quadrupole, int group_size,
                    const int32_t *group_offsets,
                    const int32_t *group_count, double theta, const ptg_criterion *criterion, double G,
                    int exact, double *out, int64_t *counters, void *ws, size_t ws_bytes, void *stream);
/* Reads the overflow flag of the last walk on this workspace (synchronises the stream):
 * PTG_ERR_OVERFLOW if a traversal stack exceeded its spill capacity, PTG_ERR_ARG if a supplied
 * group held more than group_size targets or a subtree walk's start slot was out of range. */
ptg_status ptg_walk_check(void *ws, void *stream);

/* Subtree walks: replaces the start node `no` of treewalk.PotentialWalk / AccelWalk /
 * PotentialWalk_quad / AccelWalk_quad (treewalk.py:78-277; the walk stops at subtree_end, :28-34).
 * Every target pos[perm[t]] (softening soft[perm[t]]) walks only the subtree of one element, read
 * on the device from *start_slot: that element's walk slot (the tree's slot map at the element, a
 * particle index or a node index).  A particle start contributes that particle alone; a node start
 * is tested and opened as in the whole-tree walk, so the root's slot gives ptg_walk's result and the
 * subtrees of a node's children sum to the node's own.  One tree (no forest), the FP64 walk tree with
 * the reference's branches, one target per group; the other arguments and the output as ptg_walk's
 * (ws sized by ptg_walk_workspace_bytes(nt, 1)).  A start slot outside the walk tree walks nothing
 * and is reported by ptg_walk_check as PTG_ERR_ARG. */
ptg_status ptg_walk_subtree(const ptg_walk_tree *wt, const int32_t *start_slot, const double *pos, const double *soft,
                            const int64_t *perm, int64_t nt, int fields, int quadrupole, double theta, double G,
                            double *out, int64_t *counters, void *ws, size_t ws_bytes, void *stream);

/* ptg_walk without a workspace, for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  The same
 * kernels and results as ptg_walk; the walk control block and the per-wave spill stacks live in a
 * static device array of the library (80 MiB + 4 KiB per GPU, held until the process ends: 4096
 * spill entries per wave for the 5120 waves of a 256-CU grid; measured walks need at most 1344).
 * Nothing is allocated and nothing synchronises.  err (device int32, required) receives the
 * walk's flags instead of ptg_walk_check: 0 fine, bit 0 a traversal stack exceeded the spill
 * capacity, bit 1 a supplied group held more than group_size targets.  Rejected with PTG_ERR_ARG: a
 * device whose resident grid would leave fewer than 2048 spill entries per wave.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one device (different streams) race. */
ptg_status ptg_walk_noalloc(const ptg_walk_tree *wt, const ptg_forest *forest, const double *pos, const double *soft,
                            const int64_t *perm, int targets_sorted, int64_t nt, int fields, int quadrupole,
                            int group_size, const int32_t *group_offsets, const int32_t *group_count, double theta,
                            const ptg_criterion *criterion, double G, int exact, double *out, int64_t *counters,
                            int32_t *err, void *stream);

/* ------------------------------------------------------------------------------------------
 * Learned opening criterion (no reference counterpart; it replaces the geometric test of
 * treewalk.py:22-25 by a per-node opening radius, keeping that test as the guard).  A 10-32-32-1
 * MLP with tanh hidden units maps scale-free node features and the tolerance eps (an absolute
 * acceleration, G = 1) to y = log10(r_open / size).  Parameters: PTG_MLP_NPARAM float32 values,
 * W1[10][32], b1[32], W2[32][32] (row = input unit), b2[32], W3[32], b3.
 * ------------------------------------------------------------------------------------------ */
#define PTG_MLP_NPARAM 1441
/* features[count][10] of the given node elements (indices >= n) at tolerance eps */
ptg_status ptg_node_features(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const int32_t *nodes,
                             int64_t count, double eps, float *features, void *stream);
/* Labels by direct summation: err[count][nradii] = the monopole error over 32 directions at
 * distance rho_j * size from the node's COM (rho log-spaced from rho_lo to rho_hi), reduced over
 * the directions by its maximum (stat 0) or root mean square (stat 1), in units of mass / size^2;
 * rows of nodes with more than cap particles (or massless) are set to -1.  nradii <= 64.  forest
 * (nullable) gives the particle range ends of forest trees. */
ptg_status ptg_node_errcurve(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const ptg_forest *forest,
                             const int32_t *nodes, int64_t count, int64_t cap, int nradii, double rho_lo,
                             double rho_hi, int stat, double *err, void *stream);
/* Training samples: sample i = (curve row rows[i] of node nodes[rows[i]], tolerance eps[i]) ->
 * x[i][10], label y[i] = log10 of the smallest distance (in sizes) beyond which the curve stays
 * within eps, valid[i] (0 skipped, 1 ok, 2 label clipped at rho_hi). */
ptg_status ptg_mlp_samples(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const int32_t *nodes,
                           const int32_t *rows, const double *eps, int64_t count, const double *err, int nradii,
                           double rho_lo, double rho_hi, float *x, float *y, int32_t *valid, void *stream);
/* Gradient of the loss (1/B) sum_b w_b (yhat_b - y_b)^2, w_b = under_weight where yhat_b < y_b
 * else 1, over a minibatch of B = batch samples: sample b is index[b] if index is given, else a
 * splitmix64 hash of (seed, b) modulo nsamples.  grad[PTG_MLP_NPARAM] receives the gradient and
 * grad[PTG_MLP_NPARAM] the loss (so grad holds PTG_MLP_NPARAM + 1 floats).  Deterministic. */
size_t ptg_mlp_grad_workspace_bytes(int64_t batch);
ptg_status ptg_mlp_grad(const float *params, const float *x, const float *y, int64_t nsamples, const int32_t *index,
                        uint64_t seed, int64_t batch, float under_weight, float *grad, void *ws, size_t ws_bytes,
                        void *stream);
/* One Adam step (step counts from 1) on params with moment buffers m, v. */
ptg_status ptg_mlp_adam(float *params, float *m, float *v, const float *grad, int64_t step, float lr, float beta1,
                        float beta2, float eps, void *stream);
/* y[i] = MLP(x[i]); bf16 != 0: weights and activations rounded to bfloat16, FP32 sums. */
ptg_status ptg_mlp_forward(const float *params, const float *x, int64_t count, int bf16, float *y, void *stream);
/* Fused inference: writes size * 10^(MLP(features(node, eps)) + margin) into aux[3] of every node
 * slot of an FP32 pack from ptg_pack_walk (0 for massless nodes; particle slots keep the pack's 0),
 * for PTG_CRIT_RADIUS walks.  The tolerance is eps * eps_scale[s] when eps_scale (device, e.g. the
 * RMS field from ptg_rms_rows) is given: s = 0 for one tree, the node's system for a forest (one
 * scale per system), so the tolerance never has to visit the host. */
ptg_status ptg_open_radii(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const ptg_forest *forest,
                          const ptg_walk_tree *wt, const float *params, float margin, double eps,
                          const double *eps_scale, int bf16, void *stream);
/* out[s] = sqrt(mean over the rows r of segment s of sum_{c < ncols} x[r*stride + col + c]^2) for
 * nseg segments of rows / nseg consecutive rows (no reference counterpart: the RMS field, per
 * system of a forest, that scales the criteria's tolerance); device results. */
ptg_status ptg_rms_rows(const double *x, int64_t rows, int stride, int col, int ncols, int64_t nseg, double *out,
                        void *stream);

/* ------------------------------------------------------------------------------------------
 * Column density (treewalk.py:821-1276; frontend.ColumnDensity :1321-1457).  Particles are uniform
 * spheres of radius equal to their softening; the tree is walked through its FirstSubnode /
 * NextBranch links in the reference's depth-first order.  targets (nt, 3) float64 device array;
 * err (device int32) is set to 1 if a link chain is corrupt.
 * ------------------------------------------------------------------------------------------ */
/* Ray-traced: out[j * nrays + i] = column from target j to infinity along unit ray i, exact
 * (ColumnDensityWalk_singleray :904-973 per pair; the grouped and multi-ray walks of :822-901 and
 * :1087-1194 give the same values).  rays: (nrays, 3) shared, or (nt, nrays, 3) when per_target
 * != 0 (randomized ray grids, :1061-1078).  wt: the FP64 walk pack of the same tree (ptg_pack_walk,
 * PTG_F64) for the wavefront walk (64 targets per wavefront walk the union of their paths, one
 * record load per element), or NULL for the stackless per-thread walk on the links; err 2 means
 * the tree was deeper than the wavefront walk's stack: rerun with wt = NULL.  order (nullable, used
 * with wt): (nrays, nt) permutations of the targets, one per ray, grouping 64 targets per wavefront
 * in that order (e.g. by position across the ray, so their paths overlap); results do not depend on
 * it.  counters (nullable, device int64[1], accumulated): [0] elements visited over all walks. */
ptg_status ptg_column_rays(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const ptg_walk_tree *wt,
                           const double *targets, int64_t nt, const double *rays, int64_t nrays, int per_target,
                           const int64_t *order, double *out, int32_t *err, int64_t *counters, void *stream);
/* Angular-binned (rays=None): out (nt, 6), the mean column per equal-area sky bin, with groups of
 * group_size consecutive targets sharing one traversal opened at the node's distance to the group's
 * box (ColumnDensityBinned_grouped :1197-1276; group_size 1 is ColumnDensityWalk_binned :976-1028).
 * group_size <= 64 walks the FP64 walk pack wt of the same tree (ptg_pack_walk, PTG_F64) with
 * ptg_column_workspace_bytes(group_size) of workspace (groups whose traversal passes 32768 visits are
 * finished by a second pass that deals their subtrees out to 64 wavefronts); larger groups use the
 * tree's links (wt, ws may be NULL).
 * err: 1 corrupt links, 2 a traversal stack overflow.  counters (nullable, device int64[2],
 * accumulated): [0] element-target pairs evaluated, [1] elements visited by the group traversals. */
size_t ptg_column_workspace_bytes(int64_t group_size);
/* Per-target ray grids for randomize_rays (treewalk.py:1061-1078, misc.random_rotation misc.py:6-54):
 * out (nt, nrays, 3) = rays (nrays, 3) @ R(first_seed + k) for target k, R(s) the rotation drawn from
 * the Mersenne-Twister stream seeded with integer s (x axis, then a second uniform direction made
 * orthonormal to it, z their cross product), as numpy's seeded generator draws it. */
ptg_status ptg_random_rotations(int64_t first_seed, int64_t nt, const double *rays, int64_t nrays, double *out,
                                void *stream);
ptg_status ptg_column_binned(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const ptg_walk_tree *wt,
                             const double *targets, int64_t nt, int64_t group_size, double theta, double *out,
                             int32_t *err, void *ws, size_t ws_bytes, int64_t *counters, void *stream);

/* ------------------------------------------------------------------------------------------
 * Direct summation, LDS-tiled: replaces bruteforce.py.  self_mode != 0: targets are the sources
 * (nt == ns), pairs with equal index are excluded (Potential_bruteforce_parallel :86-116,
 * Accel_bruteforce_parallel :164-206, TidalTensor_bruteforce_parallel :319-350); self_mode == 0:
 * target form, r == 0 pairs skipped (PotentialTarget_bruteforce :7-42, AccelTarget_bruteforce
 * :209-248, TidalTensorTarget_bruteforce :353-386).  out is (nt, W) float64, scaled by G.
 * ------------------------------------------------------------------------------------------ */
ptg_status ptg_direct(const double *pos_t, const double *soft_t, int64_t nt, const double *pos_s, const double *m_s,
                      const double *h_s, int64_t ns, int prec, int fields, int self_mode, double G, double *out,
                      void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PTG_H */
