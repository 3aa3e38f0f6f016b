/* voro_gpu.h -- C ABI of the MI355X (gfx950) Voronoi cell engine.
 *
 * Plain C types only: pointers, sizes, status codes.  One handle is one
 * particle container; the calls on a handle are not re-entrant (as the
 * reference's containers are not, src/v_compute.hh:103-119), and a handle is
 * used from one host thread at a time.
 *
 * Each entry point names the reference interface it stands in for
 * (paths relative to the reference checkout).
 */
#ifndef VORO_GPU_H
#define VORO_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  1..4 are the reference's exit statuses
 * (src/config.hh:103-119), returned instead of calling exit(). */
#define VGPU_OK 0
#define VGPU_FILE_ERROR 1
#define VGPU_MEMORY_ERROR 2      /* a capacity limit was exceeded */
#define VGPU_INTERNAL_ERROR 3    /* inconsistent cell (reference: fatal) */
#define VGPU_CMD_LINE_ERROR 4
#define VGPU_INVALID_ARGUMENT 5
#define VGPU_HIP_ERROR 6         /* a HIP runtime call failed */
#define VGPU_DUPLICATE 7         /* periodic put of a duplicate particle (reference: exit(1)) */
#define VGPU_NO_DEVICE 8

/* Container kinds. */
#define VGPU_CONTAINER 0                /* src/container.hh:288  container */
#define VGPU_CONTAINER_POLY 1           /* src/container.hh:510  container_poly */
#define VGPU_CONTAINER_PERIODIC 2       /* src/container_prd.hh:213 container_periodic */
#define VGPU_CONTAINER_PERIODIC_POLY 3  /* src/container_prd.hh:434 container_periodic_poly */

/* vgpu_compute_all flags. */
#define VGPU_F_NEIGHBORS 1u   /* produce neighbour lists (voronoicell_neighbor) */
#define VGPU_F_STATS 2u       /* record per-cell candidate statistics */
#define VGPU_F_LEARNED 4u     /* use the trained candidate predictor if one is loaded */
#define VGPU_F_CSR 8u         /* compact neighbour lists into loop-order CSR */
#define VGPU_F_GRAPH 16u      /* batches: capture the compute as a hipGraph once, replay it afterwards */
#define VGPU_F_GEOMETRY 32u   /* keep every cell's vertices and faces and compute the polyhedron
                                 statistics (vgpu_get_geometry); implies VGPU_F_NEIGHBORS | VGPU_F_CSR
                                 and the plain path (single containers only) */

typedef struct vgpu_container vgpu_container;

/* Creates a container.
 *   rectangular kinds: geom = {ax,bx,ay,by,az,bz}, periodic = {xp,yp,zp}
 *     (src/container.hh:290-291, container(ax_,bx_,...,nx_,ny_,nz_,xp,yp,zp,init_mem))
 *   periodic kinds:    geom = {bx,bxy,by,bxz,byz,bz}, periodic ignored
 *     (src/container_prd.hh:215-216, 434-441)
 * init_mem is accepted for signature parity and ignored: storage is sized
 * from the particle count. */
int vgpu_create(int kind, const double *geom, const int *grid, const int *periodic, int init_mem,
		vgpu_container **out);

/* Frees the container and its device memory. */
int vgpu_destroy(vgpu_container *h);

/* Adds a wall (src/container.hh:63-69 add_wall); walls cut each cell in the
 * order they were added.  Rectangular containers only, at most 64.
 *   type 0 wall_sphere(xc,yc,zc,rc,w_id)            src/wall.hh:27-30
 *   type 1 wall_plane(xc,yc,zc,ac,w_id)             src/wall.hh:44-47
 *   type 2 wall_cylinder(xc,yc,zc,xa,ya,za,rc,w_id) src/wall.hh:65-68
 *   type 3 wall_cone(xc,yc,zc,xa,ya,za,ang,w_id)    src/wall.hh:97-101
 * params: the constructor's arguments before w_id, in order (4 for the
 * sphere and the plane, 7 for the cylinder and the cone). */
int vgpu_add_wall(vgpu_container *h, int type, const double *params, int w_id);
/* vgpu_add_wall(h, 0, {xc,yc,zc,rc}, w_id). */
int vgpu_add_wall_sphere(vgpu_container *h, double xc, double yc, double zc, double rc, int w_id);
/* container_base::point_inside (src/container.cc:484-487): *inside = 1 if
 * (x,y,z) lies in the box and inside every wall (each wall's point_inside,
 * src/wall.cc). */
int vgpu_point_inside(vgpu_container *h, double x, double y, double z, int *inside);

/* Appends n particles from host memory, in call order, exactly as n calls
 * of put(id,x,y,z[,r]) would (src/container.cc:90-115,
 * src/container_prd.cc:107-123).  xyz is n x 3 doubles; r may be NULL for
 * the plain kinds and is required for the _poly kinds.  Out-of-box
 * particles of non-periodic directions are dropped, as the reference does. */
int vgpu_put(vgpu_container *h, int n, const int *ids, const double *xyz, const double *r);

/* Replaces the container's particles by n particles already in device
 * memory (same layout as vgpu_put).  Nothing is copied to the host. */
int vgpu_put_device(vgpu_container *h, int n, const int *d_ids, const double *d_xyz, const double *d_r);

/* Removes all particles (src/container.hh:332-335 clear()). */
int vgpu_clear(vgpu_container *h);

/* Bins the particles and computes every cell, storing the results
 * (src/container.cc:438-443 compute_all_cells(); the reference discards the
 * cells, this engine keeps volume, centroid, face count, existence and, with
 * VGPU_F_NEIGHBORS, neighbour ids for c_loop_all reads).  `stream` is a
 * hipStream_t (NULL for the default stream); the call returns when the
 * results are complete. */
int vgpu_compute_all(vgpu_container *h, unsigned flags, void *stream);

/* container::find_voronoi_cell / container_poly::find_voronoi_cell
 * (src/container.cc:237-297) and container_periodic(_poly)::
 * find_voronoi_cell (src/container_prd.cc:303-360) for n query points at
 * once, as n calls made one after another: found[i] = 1 with pid[i] and
 * pos[3i..3i+2] (the particle's position, shifted by whole periods to the
 * image nearest the point) for the particle whose Voronoi cell contains
 * point i, or 0 (point outside a non-periodic direction, or no particle).
 * Periodic classes: the queries create the image blocks they reach, in the
 * reference's order, on top of those the last compute created.  Bins the
 * particles if they changed since the last compute.  on_device: the four
 * arrays are device pointers.  Returns VGPU_MEMORY_ERROR (found[i] = -1) if a
 * search outgrew the largest block window (±63 blocks). */
int vgpu_find_voronoi_cells(vgpu_container *h, int n, const double *xyz, int *found, int *pid, double *pos,
			    int on_device);
/* container::compute_ghost_cell / container_poly::compute_ghost_cell
 * (src/container.hh:488-498,710-721) and container_periodic(_poly)::
 * compute_ghost_cell (src/container_prd.hh:414-420,635-645) for n points
 * (host arrays; r: radii, radical containers only): the cell a particle at
 * each point would have if it were put last in its block.  exists[i]: the
 * reference's return value (0: the point lies outside a non-periodic
 * direction, or its cell vanished); vol, cent (relative to the point),
 * nfaces, and up to max_nbr neighbour ids per point at nbr[i * max_nbr].
 * Returns VGPU_MEMORY_ERROR, with exists[i] = -1 for the points concerned and
 * every other point's results complete, for a cell over 4,096 vertices or
 * max_nbr faces, a search that comes back to the point's own block through a
 * periodic image, or (periodic classes) one that needs an image block the
 * last compute or queries did not create (the reference would create it
 * during the call and copy the point into it, with an id it never sets). */
int vgpu_compute_ghost_cells(vgpu_container *h, int n, const double *xyz, const double *r, int *exists, double *vol,
			     double *cent, int *nfaces, int *nbr, int max_nbr);
/* For each stored particle in loop order, its position in the sequence of
 * put calls (0-based, counting every particle given to vgpu_put /
 * vgpu_put_device since the last vgpu_clear, dropped ones included): the
 * record a particle_order keeps (src/c_loops.hh:35-75) turned around, for
 * c_loop_order.  Valid after a compute. */
int vgpu_get_put_index(vgpu_container *h, int *put_index);
/* The container kind (VGPU_CONTAINER ... VGPU_CONTAINER_PERIODIC_POLY). */
int vgpu_kind(vgpu_container *h, int *kind);
/* Bins the particles put so far without computing cells (what the
 * reference's put() has done by the time a loop reads them); vgpu_compute_all
 * bins by itself. */
int vgpu_bin(vgpu_container *h);
/* After binning: the particle count and, for each stored particle in loop
 * order, its id, stored position (remapped into the domain along periodic
 * directions) and radius (default_radius 0.5 for the plain kinds,
 * src/config.hh:90) -- c_loop_base::pos(pid,x,y,z,r) (src/c_loops.hh:146-153)
 * for every loop position.  Any pointer may be NULL. */
int vgpu_get_particles(vgpu_container *h, long long *n, int *id, double *pos, double *r);

/* c_loop_subset (src/c_loops.hh:218-258, src/c_loops.cc:24-150) over a
 * rectangular container's stored particles, the reference's loop state
 * machine step for step: mode 0 setup_sphere(v[0], v[1], v[2], r = v[3],
 * bounds_test), 1 setup_box(v[0..5] = xmin, xmax, ymin, ymax, zmin, zmax,
 * bounds_test), 2 setup_intbox(iv[0..5] = ai, bi, aj, bj, ak, bk).  For the
 * visited particles in the loop's order: their c_loop_all position
 * (index[t], into vgpu_get_results' arrays) and the loop's block coordinates
 * (ijk_pos[3t..3t+2] = i, j, k as the reference's loop holds them, unwrapped
 * along periodic directions); up to cap entries, *n = the loop length.  Bins
 * the particles if needed. */
int vgpu_loop_subset(vgpu_container *h, int mode, const double *v, const int *iv, int bounds_test, long long *index,
		     int *ijk_pos, long long cap, long long *n);

/* Text formats, byte for byte the reference's:
 *   print_custom(format, filename) -- container::print_custom
 *     (src/container.cc:401-432; the per-cell line of
 *     voronoicell_base::output_custom, src/cell.cc:2485-2558, every code);
 *     computes the cells (with VGPU_F_GEOMETRY if a code needs the vertices
 *     or faces);
 *   import(filename) -- container::import / container_poly::import
 *     (src/container.cc:333-376): "id x y z" (radii != 0: "id x y z r")
 *     records until the end of the file; the records read are put, and a
 *     record that does not parse returns VGPU_FILE_ERROR (the reference exits);
 *   draw_particles(filename, pov) -- draw_particles / draw_particles_pov
 *     (src/container.hh:329-372, container_poly's 551-596);
 *   draw_cells(filename, pov) -- draw_cells_gnuplot / draw_cells_pov
 *     (src/container.hh:379-418; voronoicell_base::draw_gnuplot and
 *     draw_pov, src/cell.cc:1940-1987). */
int vgpu_print_custom(vgpu_container *h, const char *format, const char *filename);
int vgpu_import(vgpu_container *h, const char *filename, int radii, long long *n_read);
int vgpu_draw_particles(vgpu_container *h, const char *filename, int pov);
int vgpu_draw_cells(vgpu_container *h, const char *filename, int pov);

/* Number of particles stored (the c_loop_all length). */
int vgpu_num_cells(vgpu_container *h, long long *n);

/* Copies results to host arrays, in c_loop_all order (blocks by ascending
 * index, then storage order; src/c_loops.hh:165-209).  Any pointer may be
 * NULL.  nbr_off has n+1 entries; nbr has nbr_cap entries and receives
 * neighbour ids in the order voronoicell_neighbor::neighbors() gives them
 * (src/cell.cc:2660-2678).  *nbr_total receives the total id count even
 * when it exceeds nbr_cap. */
int vgpu_get_results(vgpu_container *h, int *id, int *ijk, int *q, int *exists, double *pos, double *vol,
		     double *centroid, int *nfaces, long long *nbr_off, int *nbr, long long nbr_cap,
		     long long *nbr_total);

/* Per-cell statistics recorded with VGPU_F_STATS: nplane calls, worklist
 * exit index (64 = reached the block search), and the tier that produced
 * the cell (0 first tier: one cell per wavefront, order-3 per lane, or with
 * VGPU_F_LEARNED the list-driven one; 1 order-3 over the first tier's
 * leftovers; 2 wide; 3 huge).  Loop order. */
int vgpu_get_stats(vgpu_container *h, int *n_planes, int *wl_exit, int *tier);

/* Polyhedron statistics of every cell, after a vgpu_compute_all with
 * VGPU_F_GEOMETRY, in loop order, each equal bit for bit to what the
 * reference's voronoicell_neighbor function returns for that cell
 * (src/cell.cc): vertices() (local coordinates, 3 per vertex; 1791-2216),
 * vertex_orders() (2191), the edge table ed[i][j] and its back pointers
 * ed[i][nu[i]+j] (vertex by vertex, nu[i] slots each), face_areas() (1791),
 * face_perimeters() (2251), normals() (2087, 3 per face), face_orders()
 * (2309), face_vertices() (2282, per face: the count, then the vertices),
 * surface_area() (1822), total_edge_distance() (1921), number_of_edges()
 * (2468), max_radius_squared() (1907).  Per-vertex, per-edge-slot and
 * face-vertex arrays are indexed by the offsets returned here (n+1 each);
 * per-face arrays by the neighbour offsets of vgpu_get_results (one
 * neighbour per face, same order).  Cells that do not exist have none.
 * Any pointer may be NULL; sizes from vgpu_geometry_sizes. */
int vgpu_geometry_sizes(vgpu_container *h, long long *n_vertices, long long *n_edge_slots, long long *n_faces,
			long long *n_face_vertex_entries);
int vgpu_get_geometry(vgpu_container *h, long long *vert_off, double *vertices, int *vertex_orders, long long *edge_off,
		      int *edges, int *edge_rel, double *face_areas, double *face_perimeters, double *normals,
		      int *face_orders, long long *fv_off, int *face_vertices, double *surface_area,
		      double *total_edge_distance, int *number_of_edges, double *max_radius_squared);

/* Sum of cell volumes in loop order (src/container.cc:451-456
 * sum_cell_volumes()). */
int vgpu_sum_volumes(vgpu_container *h, double *out);

/* Device pointers to the loop-order results (valid until the next put or
 * compute).  Any pointer may be NULL. */
int vgpu_device_results(vgpu_container *h, const double **vol, const double **centroid, const int **nfaces,
			const int **status, const int **nbr_start, const int **nbr);

/* Tuning of the fast-tier cell kernel: the number of persistent
 * wavefronts (0 = 4 x occupancy x compute units) and the register budget it
 * was compiled for, in waves per SIMD (2, 3 or 4; 0 = keep).  Results do not
 * depend on either. */
int vgpu_set_launch(vgpu_container *h, int fast_waves, int fast_occupancy);

/* The first tier of the plain path of rectangular containers: 0 (default)
 * computes one cell per wavefront, vertex v in lane v; 1 one order-3 cell per
 * lane.  Cells the first tier cannot finish go on to the per-lane tiers, so
 * results do not depend on the choice.  waves: persistent wavefronts of the
 * wavefront tier (0 = 4 x occupancy x compute units); occupancy: its
 * register budget in waves per SIMD (2, 3, 4, 5, 6 or 8, default 5, the
 * learned path's list kernel and periodic containers 4, 5 or 6; 0 =
 * keep). */
int vgpu_set_first_tier(vgpu_container *h, int tier, int waves, int occupancy);

/* Autotuner for one container (no counterpart in the reference): times
 * vgpu_compute_all(h, flags) `reps` times for each launch configuration and
 * keeps the fastest on the handle (of configurations within 1% of the
 * fastest, the one with the fewest waves per SIMD: it spills least to
 * memory).  The configurations are the first tier's
 * cells per wavefront (one wavefront cell, vertex tables in registers, at 3,
 * 4, 5 or 6 waves per SIMD -- periodic containers: 4, 5 or 6 -- or 64 order-3 cells
 * per wavefront, one per lane, at 2, 3 or 4) and, with VGPU_F_LEARNED, the
 * candidate-list length (the predictor's depth margin, -2 .. +4 worklist
 * steps around the trained one, at the best register budget).  Results do
 * not depend on the choice.
 * trials[0] is the chosen configuration, then the others in the order tried
 * (up to max_trials); *n_trials = configurations tried.  flags may not hold
 * VGPU_F_GEOMETRY. */
typedef struct {
	int first_tier;       /* 0 one cell per wavefront, 1 one cell per lane */
	int occupancy;        /* register budget, waves per SIMD */
	int waves;            /* persistent wavefronts */
	float depth_offset;   /* learned path: added to the trained depth margin */
	double ms;            /* fastest of the timed computes */
} vgpu_tune_trial;
int vgpu_autotune(vgpu_container *h, unsigned flags, int reps, vgpu_tune_trial *trials, int max_trials,
		  int *n_trials);

/* Time of the last vgpu_compute_all, split by stage, in milliseconds, from
 * HIP events: [0] binning, [1] fast tier, [2] wide+huge tiers, [3] outputs. */
int vgpu_last_timings(vgpu_container *h, double *ms4);

/* ---- Learned candidate-depth predictor (no counterpart in the reference) ----
 * A small perceptron (28 inputs: the particle counts of the 3x3x3 blocks
 * around a block divided by the mean count, and in radical mode the radius
 * divided by the largest one; two hidden layers of 32, ReLU; FP32) predicts
 * the worklist depth at which the security-radius test stops a cell.  With
 * VGPU_F_LEARNED, vgpu_compute_all gathers each cell's candidates up to the
 * predicted depth into a list and cuts from the list; a cell whose list ends
 * before the radius test stops it is recomputed by the plain path, so the
 * results are identical with or without the predictor.  Training runs on
 * the GPU (forward, backward and Adam kernels). */
typedef struct vgpu_predictor vgpu_predictor;

/* with_radius: predict per particle with the radius as input 28 (for the
 * _poly kinds); otherwise per block. */
int vgpu_predictor_create(int with_radius, vgpu_predictor **out);
int vgpu_predictor_destroy(vgpu_predictor *p);
/* Appends training samples from the last vgpu_compute_all of h, which must
 * have been run with VGPU_F_STATS (the plain path's own statistics are the
 * labels). */
int vgpu_predictor_add_samples(vgpu_predictor *p, vgpu_container *h);
int vgpu_predictor_clear_samples(vgpu_predictor *p);
int vgpu_predictor_num_samples(vgpu_predictor *p, long long *n);
/* Full-batch Adam for `epochs` steps at learning rate lr (parameters start
 * from a seeded initialisation unless set before), then sets the depth
 * margin so that a fraction `fallback` of the samples would be
 * under-predicted.  *loss receives the final mean squared error. */
int vgpu_predictor_train(vgpu_predictor *p, int epochs, float lr, unsigned seed, float fallback, float *loss);
/* Reads (get) and/or replaces (set) the n = vgpu_mlp_num_params() weights. */
int vgpu_predictor_params(vgpu_predictor *p, float *get, const float *set, int n);
int vgpu_predictor_margin(vgpu_predictor *p, float *get, const float *set);
/* Reads and/or replaces the depth from which a cell is given no list (its
 * predicted depth, margin included, says it reaches the block search, which
 * no list covers; it takes the plain path at once).  Default 64; 65 or more
 * gives every cell a list of depth at most 63. */
int vgpu_predictor_route(vgpu_predictor *p, int *get, const int *set);
/* Uses p (NULL: none) for VGPU_F_LEARNED computations of h. */
int vgpu_set_predictor(vgpu_container *h, vgpu_predictor *p);
/* Cells of the last learned computation whose list ended before the cell
 * was finished (they fell back to the plain path), the list records
 * gathered (one per worklist step, including batch padding), the time from the end of binning to
 * the end of list gathering (features, prediction, gathering; included in
 * vgpu_last_timings' first-tier figure), and the cells given no list because
 * the block search was predicted (they took the plain path at once).  Any
 * pointer may be NULL. */
int vgpu_learned_stats(vgpu_container *h, long long *fallbacks, long long *list_entries, double *gather_ms,
		       long long *routed);
/* The network's kernels on host data, for checking them: loss
 * (1/n) sum (y - t)^2 and its gradient, and forward outputs (n x 2). */
int vgpu_mlp_num_params(void);
int vgpu_mlp_loss_grad(const float *params, const float *X, const float *T, int n, float *loss, float *grad);
int vgpu_mlp_forward(const float *params, const float *X, int n, float *Y);
/* vgpu_mlp_loss_grad without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight training loops.  params (vgpu_mlp_num_params()
 * floats), X (n x 28), T (n x 2), grad (vgpu_mlp_num_params() floats) and
 * loss (1 float, may be NULL) are DEVICE pointers; stream is a hipStream_t
 * (NULL: the default stream); the call returns when the results are
 * written.  The same kernels and summation order as vgpu_mlp_loss_grad, so
 * the results are bit-identical to it.  Its scratch is a fixed device array
 * of the library (17.6 MB per GPU, held until the process ends) sized for
 * at most VGPU_MLP_NOALLOC_MAX_SAMPLES samples: larger n returns
 * VGPU_INVALID_ARGUMENT.  All calls on one GPU share that array, so calls
 * that overlap on one GPU (other host threads, other streams) race. */
#define VGPU_MLP_NOALLOC_MAX_SAMPLES 32768
int vgpu_mlp_loss_grad_noalloc(const float *params, const float *X, const float *T, int n, float *loss, float *grad,
			       void *stream);

/* ---- Batches of containers (no counterpart in the reference) ----
 * Many independent rectangular containers -- each created, walled and
 * filled through the calls above, exactly as the reference's program would
 * set up one container object per system -- computed together: one binning
 * pass and one persistent work-queue cell kernel over all of them, in place
 * of a compute_all_cells() (src/container.cc:438-443) per container.  Every
 * container's results are those vgpu_compute_all would give it.  All
 * containers of a batch must be of one kind (container or container_poly). */
typedef struct vgpu_batch vgpu_batch;

/* Copies the containers' geometry, walls and particles (as put so far) into
 * a new batch; the handles are not kept. */
int vgpu_batch_create(vgpu_container *const *hs, int count, vgpu_batch **out);
int vgpu_batch_destroy(vgpu_batch *b);
/* Computes every container (flags as vgpu_compute_all; VGPU_F_GRAPH
 * captures the whole device sequence as a hipGraph on the first call and
 * replays it on later calls with the same flags; VGPU_F_LEARNED needs a
 * predictor from vgpu_batch_set_predictor and works with VGPU_F_GRAPH;
 * VGPU_F_GEOMETRY is not available for batches). */
int vgpu_batch_compute(vgpu_batch *b, unsigned flags, void *stream);
/* Uses p (a block predictor, not one that takes the radius; NULL: none) for
 * VGPU_F_LEARNED computations of b.  Drops a captured graph. */
int vgpu_batch_set_predictor(vgpu_batch *b, vgpu_predictor *p);
/* The last learned computation of b: cells whose list ended first, cells
 * given no list (block search predicted, or their list did not fit), list
 * records written, and the records every list needed (more than the
 * capacity: the next computation gets room for all of them).  Any pointer
 * may be NULL. */
int vgpu_batch_learned_stats(vgpu_batch *b, long long *fallbacks, long long *routed, long long *list_entries,
			     long long *list_need);
/* Reads and/or replaces the candidate-list capacity (records of 32 bytes)
 * of b's learned computations.  Lists beyond it are not written; their cells
 * take the plain path with identical results.  Drops a captured graph. */
int vgpu_batch_list_capacity(vgpu_batch *b, long long *get, const long long *set);
/* Training samples from the last vgpu_batch_compute of b (run with
 * VGPU_F_STATS): per block of every container, as vgpu_predictor_add_samples
 * takes them from one container (a block predictor only). */
int vgpu_predictor_add_batch_samples(vgpu_predictor *p, vgpu_batch *b);
/* Replaces all particle positions (same count and order as at creation:
 * n x 3 doubles, host or device memory). */
int vgpu_batch_set_positions(vgpu_batch *b, const double *xyz, int on_device);
/* Container `index`'s cell count and results, as vgpu_num_cells and
 * vgpu_get_results give them for a single container. */
int vgpu_batch_num_cells(vgpu_batch *b, int index, long long *n);
int vgpu_batch_get_results(vgpu_batch *b, int index, int *id, int *ijk, int *q, int *exists, double *pos, double *vol,
			   double *centroid, int *nfaces, long long *nbr_off, int *nbr, long long nbr_cap,
			   long long *nbr_total);
/* Per-container sum_cell_volumes() (out has one entry per container). */
int vgpu_batch_sum_volumes(vgpu_batch *b, double *out);
/* Launch configuration {persistent wavefronts of the first tier (0 = as
 * many as fit), first tier (2..4: one cell per lane at that many waves per
 * SIMD, default 2; 0, 5, 6: one cell per wavefront at 4, 5 or 6 waves per
 * SIMD, which the learned path's list tier also takes), cells per queue
 * entry (1..64, default 64)}; get and/or set.  Results do not depend on it. */
int vgpu_batch_config(vgpu_batch *b, int *get3, const int *set3);
/* Times every configuration on this batch (reps computes each) and keeps
 * the fastest; best3 / best_ms receive it and its time. */
int vgpu_batch_autotune(vgpu_batch *b, unsigned flags, int reps, int *best3, double *best_ms);
/* Milliseconds of the last compute: [0] binning and queue set-up, [1] first
 * tier, [2] retry tiers and neighbour compaction (graph replays: the total
 * in [1]). */
int vgpu_batch_last_timings(vgpu_batch *b, double *ms3);

/* Human-readable text of a status code. */
const char *vgpu_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
