/*
 * illcrawl_psg.h -- C ABI of the MI355X-native projective SPH smoothing path.
 *
 * The entry points are the device-side boundary that illcrawl's projective
 * smoothing backend stack calls.  Each declaration names the reference
 * interface it replaces (file:line in the illcrawl sources).
 *
 * Conventions
 *  - Every function returns an ilc_status; on failure ilc_last_error()
 *    returns a thread-local message.  Nothing throws across the ABI.
 *  - Pointers named *_dev are device (HBM) pointers; they are borrowed for
 *    the duration of the call (and of the work it enqueues on `stream`).
 *  - `stream` is a hipStream_t (NULL = the default stream).  Work is
 *    enqueued asynchronously unless the description says otherwise.
 *  - A handle is used by one host thread at a time.
 *  - Layouts follow the reference: particles are float4 (x, y, z, w)
 *    (particle_grid.hpp:43-48); images are row-major with index ix + nx*iy
 *    (camera.cl:37, multi_array.hpp:503-536).
 */
#ifndef ILLCRAWL_PSG_H
#define ILLCRAWL_PSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  ILC_OK = 0,
  ILC_EINVAL = 1,    /* invalid argument or precondition */
  ILC_ENOMEM = 2,    /* device or host allocation failed */
  ILC_EHIP = 3,      /* HIP runtime error */
  ILC_ENOTREADY = 4, /* call order violated (e.g. run before setup) */
  ILC_ERANGE = 5     /* data outside what the grid can index */
} ilc_status;

/* Culling semantics of the gather. */
typedef enum {
  /* The reference's: per point, cells within max h of the point's own
   * cell + the cell half-diagonal (capped by the overall max h), from the
   * unclamped cell key (smoothing_particle_grid.cl:54-100). */
  ILC_MODE_REFERENCE = 0,
  /* Every particle whose support covers the point. */
  ILC_MODE_EXACT = 1
} ilc_mode;

/* Projected smoothing kernel. */
typedef enum {
  /* 2*1.04445/h^3 * 1.066667 (h^2-r^2)^2.5 / h^4: the reference's projected
   * quartic including its factor 2 (smoothing_kernel.cl:46-61). */
  ILC_KERNEL_QUARTIC_REF = 0,
  /* the same shape normalised to unit integral: 7/(2 pi h^2) (1-q^2)^2.5 */
  ILC_KERNEL_QUARTIC = 1,
  /* projected M4 cubic spline, support 2h (smoothing_kernel.cl:64-84 in 3-D),
   * tabulated on q^2 */
  ILC_KERNEL_CUBIC = 2
} ilc_kernel;

typedef struct ilc_psg ilc_psg;

typedef struct {
  int device;                        /* HIP device ordinal */
  uint64_t target_particles_per_cell; /* 0 -> the reference's default: 1000
                                        for the projective grid
                                        (projective_smoothing_backend_grid.cpp:29),
                                        64 for the volumetric grid
                                        (dm_reconstruction_backend_grid.hpp:39) */
  int mode;                          /* ilc_mode */
  int kernel;                        /* ilc_kernel */
  int tile_w, tile_h;                /* gather tile in points; 0 -> autotuned default */
} ilc_psg_config;

typedef struct {
  int32_t num_cells[3];   /* particle_grid::get_num_grid_cells (particle_grid.cpp:126-129) */
  float min_corner[3];    /* particle_grid::get_grid_min_corner (:121-124) */
  float cell_size[3];     /* particle_grid::get_grid_cell_sizes (:131-134) */
  float max_h;            /* smoothing_particle_grid::get_overall_max_smoothing_length
                             (smoothing_particle_grid.cpp:68-72) */
  uint64_t num_particles; /* particle_grid::get_num_particles (:342-346) */
} ilc_grid_info;

/* Camera of illcrawl's camera class (camera.hpp:35-56) plus the
 * projective wrapper's integration depth
 * (projective_smoothing_reconstruction.hpp:61-63). */
typedef struct {
  double position[3];
  double look_at[3];      /* unit vector */
  double roll;            /* radians */
  double width;           /* screen width; pixel size = width / nx */
  uint32_t nx, ny;
  double depth;           /* particles with camera z outside [0, depth] get w = 0 */
} ilc_camera;

/* ---------------------------------------------------------------- errors */

/* Thread-local message of the last failure (qcl::qcl_error /
 * std::runtime_error text in the reference, qcl.hpp:79-118). */
const char* ilc_last_error(void);
/* Version of this ABI. */
int ilc_abi_version(void);

/* ---------------------------------------- projective_smoothing_grid backend */

/* dm::projective_smoothing_grid constructor
 * (projective_smoothing_backend_grid.cpp:27-31; dm::grid ctor,
 * dm_reconstruction_backend_grid.cpp:26-37). */
ilc_status ilc_psg_create(const ilc_psg_config* cfg, ilc_psg** out);
ilc_status ilc_psg_destroy(ilc_psg* h);

/* reconstruction_backend::get_backend_name (reconstruction_backend.hpp:43;
 * dm_reconstruction_backend_grid.cpp:51-55).  Static string. */
const char* ilc_psg_backend_name(const ilc_psg* h);

/* reconstruction_backend::init_backend (reconstruction_backend.hpp:45;
 * dm_reconstruction_backend_grid.cpp:57-61): largest particle block a
 * later setup may pass. */
ilc_status ilc_psg_init_backend(ilc_psg* h, uint64_t blocksize);

/* projective_smoothing_backend::setup_projected_particles
 * (projective_smoothing_backend.hpp:35-42;
 * projective_smoothing_backend_grid.cpp:33-45, which continues into
 * dm::grid::setup_particles, dm_reconstruction_backend_grid.cpp:65-85, and
 * the smoothing_particle_grid / particle_grid constructors).
 * particles4_dev: n float4 (x, y, z, w) in the camera frame; hsml_dev: n
 * smoothing lengths (the backend's one additional dataset).  Builds the
 * reference's grid on the device.  Synchronises `stream` once (the grid
 * size depends on the particle bounding box, as in the reference's host
 * code, particle_grid.cpp:43-100). */
ilc_status ilc_psg_setup_projected_particles(ilc_psg* h, const float* particles4_dev,
                                             uint64_t n, const float* hsml_dev,
                                             void* stream);

/* reconstruction_backend::setup_evaluation_points
 * (reconstruction_backend.hpp:50-51; dm_reconstruction_backend_grid.cpp:87-107):
 * points4_dev: n float4 points (camera frame), kept by pointer until the next
 * call.  layout_nx: row length of the point array when it is an image
 * (tiles are formed on it); 0 = treat the points as one row.  Zeroes the
 * FP64 result accumulator. */
ilc_status ilc_psg_setup_evaluation_points(ilc_psg* h, const float* points4_dev, uint64_t n,
                                           uint32_t layout_nx, void* stream);

/* reconstruction_backend::run (reconstruction_backend.hpp:53;
 * dm_reconstruction_backend_grid.cpp:109-134): adds the current particle
 * block's contribution to every point (results accumulate until the next
 * setup_evaluation_points). */
ilc_status ilc_psg_run(ilc_psg* h, void* stream);

/* reconstruction_backend::retrieve_results (reconstruction_backend.hpp:41)
 * fused with the projective wrapper's dA scaling
 * (projective_smoothing_reconstruction.cpp:49-84, math.cl:94-102):
 * out_dev[i] = float(acc[i] * scale).  Unlike the reference, the stored
 * accumulator is not modified, so retrieving twice gives the same image. */
ilc_status ilc_psg_retrieve_results(ilc_psg* h, float scale, float* out_dev, void* stream);

/* Adjoint of run + retrieve_results(scale); the reference has no
 * counterpart (it renders only).  For point values
 *   value[i] = scale * sum_j w_j K s_ij^2.5 / h_j^2,  s = 1 - r_ij^2 / h_j^2,
 * with the projected particles of setup_projected_particles (x, y, -, w) and
 * their h, and dl_dvalue_dev[i] = dL/dvalue[i] (npts floats, device), writes
 * per particle, in input order (n floats each, device):
 *   dl_dw = dL/dw, dl_dh = dL/dh, dl_dx = dL/dx, dl_dy = dL/dy
 * (x, y in the projected frame).  Weights are recomputed, not stored.
 * Exact mode only (ILC_EINVAL otherwise): the reference-mode search box
 * makes the image a non-smooth function of the particles.  For the cubic
 * kernel, value[i] = scale * sum_j w_j F(r^2/h^2) / h^2 with F and dF/du read
 * from 4096-node tables, as the forward reads F.  Gradients are FP32 sums whose order varies from call to call
 * (atomic adds), so repeated calls agree to rounding, not bit for bit. */
ilc_status ilc_psg_backward(ilc_psg* h, const float* dl_dvalue_dev, float scale, float* dl_dw_dev,
                            float* dl_dh_dev, float* dl_dx_dev, float* dl_dy_dev, void* stream);
/* The FP64 accumulator itself (n doubles). */
ilc_status ilc_psg_retrieve_results_f64(ilc_psg* h, double* out_dev, void* stream);

/* Grid parameters of the last setup (particle_grid getters). */
ilc_status ilc_psg_grid_info(const ilc_psg* h, ilc_grid_info* out);

/* The grid in the reference's layout, for parity checks: cells2_dev
 * (2*ncells int32: begin, end; (0,0) when empty; particle_grid.cpp:104-113,
 * particle_grid.cl:114-190), maxh_dev (ncells floats,
 * grid3d_determine_max_per_cell), perm_dev (n uint32: sorted position ->
 * input index, the sorted_to_unsorted map of particle_grid.hpp:77-93, in
 * the reference's within-cell order).  Any pointer may be NULL.  Returning
 * perm synchronises the stream: the gather orders a cell's particles by
 * descending h, and the reference's order is rebuilt with a sort. */
ilc_status ilc_psg_export_grid(const ilc_psg* h, int32_t* cells2_dev, float* maxh_dev,
                               uint32_t* perm_dev, void* stream);

/* Number of (point, particle) pairs with nonzero kernel coefficient
 * (w != 0) and distance below the kernel support, per point written to
 * counts_dev (n uint64, may be NULL) and in total to *total (synchronises
 * the stream).  Independent of the culling
 * mode: this is the unit of work "particle-kernel evaluation".  No
 * reference counterpart (measurement only). */
ilc_status ilc_psg_count_evals(ilc_psg* h, uint64_t* counts_dev, uint64_t* total,
                               void* stream);

/* -------------------------------------------- camera / projection wrapper */

/* Camera basis, pixel size and lower-left screen corner in double, as
 * camera::update_basis_vectors / update_min_position compute them
 * (camera.cpp:124-155). */
ilc_status ilc_camera_basis(const ilc_camera* cam, double b0[3], double b1[3],
                            double min_position[3], double* pixel_size);

/* The FP32 projection ilc_project_particles applies, x' = minv (x - center)
 * (projective_smoothing_reconstruction.cpp:175-193, math.cl:45-91): minv is
 * row-major.  Its first two rows are the Jacobian that carries projected-frame
 * gradients back to world positions. */
ilc_status ilc_camera_projection(const ilc_camera* cam, float minv[9], float center[3]);

/* projective_smoothing::project_particles (projective_smoothing_reconstruction.cpp:175-205;
 * projection.cl:56-95): out4 = camera-frame float4 with z = 0 and w = 0 for
 * particles outside [0, depth].  in4_dev may equal out4_dev. */
ilc_status ilc_project_particles(const ilc_camera* cam, const float* in4_dev, uint64_t n,
                                 float* out4_dev, void* stream);

/* camera::generate_pixel_coordinates + projective_smoothing::project_evaluation_points
 * (camera.cpp:157-189, camera.cl:23-41, projection.cl:97-122): nx*ny float4
 * pixel-corner points in the camera frame. */
ilc_status ilc_camera_pixel_points(const ilc_camera* cam, float* out4_dev, void* stream);

/* camera::generate_pixel_coordinates alone (camera.cpp:157-189,
 * camera.cl:23-41): nx*ny world-frame pixel-corner float4 (w = 0). */
ilc_status ilc_camera_generate_pixel_coordinates(const ilc_camera* cam, float* out4_dev,
                                                 void* stream);

/* projective_smoothing::project_evaluation_points (projective_smoothing_reconstruction.cpp:144-172,
 * projection.cl:97-122): n world-frame float4 points into the camera frame, z
 * and w kept.  Unlike the reference (which overwrites the caller's buffer),
 * the result goes to out4_dev; in4_dev may equal out4_dev. */
ilc_status ilc_project_evaluation_points(const ilc_camera* cam, const float* in4_dev, uint64_t n,
                                         float* out4_dev, void* stream);

/* vector_scale (math.cl:94-102), as projective_smoothing::retrieve_results
 * applies the quantity's effective dA (projective_smoothing_reconstruction.cpp:49-84):
 * x[i] *= alpha in FP32, in place. */
ilc_status ilc_vector_scale(float* x_dev, uint64_t n, float alpha, void* stream);

/* ------------------------------------------------------- crawler block body */

/* volume_cutout (reconstructing_data_crawler.hpp:49-76): an axis-aligned
 * render volume, the period of the simulation box, and the crawler's
 * additional border region (reconstructing_data_crawler.cpp:188-193). */
typedef struct {
  double center[3];
  double extent[3];
  double periodic_wraparound[3];
  double border;
} ilc_volume;

/* The per-row body of reconstructing_data_crawler::run
 * (reconstructing_data_crawler.cpp:180-213): row i's float coordinates
 * xyz_dev[3i..3i+2] are moved by one period towards the volume grown by the
 * border (coordinate_system.hpp:24-47, in double), and the row is kept if the
 * volume contains the result (reconstructing_data_crawler.cpp:59-71).  Kept
 * rows are written in input order: particles4_dev gets float4 (x, y, z, 0),
 * each columns_out_dev[k] gets columns_dev[k][i] (ncolumns <= 8: the
 * backend's additional datasets and the quantity inputs), index_out_dev (may
 * be NULL) the input row.  *n_out = rows kept.  Output buffers hold n
 * entries.  Synchronises `stream`. */
ilc_status ilc_crawl_block(const ilc_volume* vol, const float* xyz_dev, uint64_t n,
                           const float* const* columns_dev, int32_t ncolumns,
                           float* particles4_dev, float* const* columns_out_dev,
                           uint64_t* index_out_dev, uint64_t* n_out, void* stream);

/* ilc_crawl_block without any allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and synchronisation as
 * ilc_crawl_block; its per-tile counts and their scan live in a static
 * device array of 1 MiB per GPU, held until the process ends, so blocks of
 * more than ILC_CRAWL_STATIC_MAX_ROWS rows are refused (ILC_ERANGE).  Every
 * call on one GPU uses that GPU's array: calls that overlap on one GPU (other
 * host threads, other streams) race. */
#define ILC_CRAWL_STATIC_MAX_ROWS (1ull << 29)
ilc_status ilc_crawl_block_static(const ilc_volume* vol, const float* xyz_dev, uint64_t n,
                                  const float* const* columns_dev, int32_t ncolumns,
                                  float* particles4_dev, float* const* columns_out_dev,
                                  uint64_t* index_out_dev, uint64_t* n_out, void* stream);

/* ------------------------------------------- volumetric smoothing (dm::grid) */

/* reconstruction_backends::dm::grid (dm_reconstruction_backend_grid.hpp:32-60,
 * dm_reconstruction_backend_grid.cpp:57-134): particles keep their 3-D
 * position; every evaluation point (x, y, z) gets the sum of
 * w * quartic_polynomial3d(|x_p - x|, h) (smoothing_kernel.cl:27-37) over the
 * particles of the cells in its search box (smoothing_particle_grid.cl:31-127
 * without SPG_PROJECTION), e.g. a volumetric slice (volumetric_slice.cpp).
 * The configuration is ilc_psg_config with a quartic kernel
 * (ILC_KERNEL_QUARTIC_REF: the reference's 1.04445f normalisation;
 * ILC_KERNEL_QUARTIC: 105 / (32 pi)); target 0 means the reference's 64
 * particles per cell; reference mode keeps the reference's
 * search box, exact mode sums every particle whose support contains the
 * point.  Results accumulate over run() calls until the next
 * setup_evaluation_points, as for ilc_psg. */
typedef struct ilc_grid3d ilc_grid3d;
ilc_status ilc_grid3d_create(const ilc_psg_config* cfg, ilc_grid3d** out);
ilc_status ilc_grid3d_destroy(ilc_grid3d* h);
ilc_status ilc_grid3d_init_backend(ilc_grid3d* h, uint64_t blocksize);
/* dm::grid::setup_particles (dm_reconstruction_backend_grid.cpp:65-85):
 * particles4_dev (x, y, z, w) world frame, hsml_dev one h per particle;
 * builds the 3-D grid (particle_grid.cpp:43-222), synchronises. */
ilc_status ilc_grid3d_setup_particles(ilc_grid3d* h, const float* particles4_dev, uint64_t n,
                                      const float* hsml_dev, void* stream);
/* points4_dev (x, y, z, -); layout_nx: width of the point image the tiles
 * follow (0: 16). */
ilc_status ilc_grid3d_setup_evaluation_points(ilc_grid3d* h, const float* points4_dev,
                                              uint64_t n, uint32_t layout_nx, void* stream);
ilc_status ilc_grid3d_run(ilc_grid3d* h, void* stream);
ilc_status ilc_grid3d_retrieve_results(ilc_grid3d* h, float scale, float* out_dev, void* stream);
ilc_status ilc_grid3d_retrieve_results_f64(ilc_grid3d* h, double* out_dev, void* stream);
ilc_status ilc_grid3d_grid_info(const ilc_grid3d* h, ilc_grid_info* out);
/* Exact particle-kernel evaluation count: per point (counts_dev, n points,
 * optional) the particles with w != 0 whose support contains it, and the
 * total, as ilc_psg_count_evals.  No reference counterpart (measurement
 * only); synchronises the stream. */
ilc_status ilc_grid3d_count_evals(ilc_grid3d* h, uint64_t* counts_dev, uint64_t* total,
                                  void* stream);

/* Combine particles with their quantity value
 * (reconstructing_data_crawler.cpp:233-239): particles4_dev[4i+3] = w_dev[i]. */
ilc_status ilc_particles_set_w(float* particles4_dev, const float* w_dev, uint64_t n,
                               void* stream);

/* ------------------------------------------------- quantity transformation */

/* The reference's X-ray quantities (quantities.cl:39-239, 354-430): thermal
 * bremsstrahlung of a hydrogen plasma in keV/s/kpc^3/m^2 from density rho
 * (M_sun/kpc^3), specific internal energy u ((km/s)^2, temperature as
 * get_temperature) and electron abundance xe, at redshift z and luminosity
 * distance lum_dist (kpc).  gaunt_host: the van Hoof (2014) thermally averaged
 * free-free Gaunt table as the reference loads it (gaunt_factor.cpp:30-41):
 * gny rows of constant ln(u) from -16, gnx columns of constant ln(gamma^2)
 * from -6, spacing 0.2, row-major; sampled like read_imagef with a linear,
 * clamp-to-border (0) sampler.  exp, log, pow and rsqrt are correctly
 * rounded FP32 (the reference leaves them to the OpenCL runtime).
 * Synchronous (copies the table to the device and waits). */

/* xray_flux: the band [e_min, e_max] keV (observer frame, times 1+z)
 * integrated with the reference's adaptive Runge-Kutta-Fehlberg scheme
 * (integration.cl:38-131; relative step tolerance 1/num_samples).
 * steps_dev (may be NULL): RKF steps taken per particle (at most 65536). */
ilc_status ilc_quantity_xray_flux(const float* rho_dev, const float* u_dev, const float* xe_dev,
                                  uint64_t n, const float* gaunt_host, int32_t gnx, int32_t gny,
                                  float z, float lum_dist, float e_min, float e_max,
                                  int32_t num_samples, float* w_dev, int32_t* steps_dev,
                                  void* stream);

/* xray_spectral_flux: the emission in one energy bin of width dE at E keV. */
ilc_status ilc_quantity_xray_spectral_flux(const float* rho_dev, const float* u_dev,
                                           const float* xe_dev, uint64_t n,
                                           const float* gaunt_host, int32_t gnx, int32_t gny,
                                           float z, float lum_dist, float e, float de,
                                           float* w_dev, void* stream);

/* quantity_transformation with unprocessed_quantity / constant_quantity
 * (quantity.cpp:84-131, quantities.cl:701-726): w[i] = float(scale * in[i])
 * (in_dev == NULL: w[i] = float(scale)). */
ilc_status ilc_quantity_scaled(const float* in_dev, double scale, uint64_t n, float* w_dev,
                               void* stream);

/* quantity_transformation with stellar_mass (quantity.hpp:950-1017,
 * quantities.cl:728-750): w[i] = mass[i] where the formation time is > 0 (a
 * star), 0 where it is not (a wind cell); the inputs as the crawler scaled
 * them (formation time x 1, mass x the mass-per-volume unit factor). */
ilc_status ilc_quantity_stellar_mass(const float* formation_time_dev, const float* mass_dev,
                                     uint64_t n, float* w_dev, void* stream);

/* Bolometric free-free emissivity w = n_e n_i sqrt(T) g(T), n_e = n_i =
 * 4.046263 rho x_e (the ion density convention of quantities.cl:112), g read
 * from a 1-D table (ntab floats, node i at x0 + i*dx on ln T) with the
 * reference's image sampler semantics (tabulated_function.cl:31-44: linear
 * filter, half-texel centres, zero border).  The table is staged in LDS. */
ilc_status ilc_quantity_xray_bolometric(const float* rho_dev, const float* temp_dev,
                                        const float* xe_dev, uint64_t n,
                                        const float* table_host, int32_t ntab, float x0,
                                        float dx, float* w_dev, void* stream);


/* ------------------------------------------------- batched multi-view path */

/* Many projections of one resident snapshot in one persistent gather launch,
 * optionally replayed from a hipGraph.  Replaces, on one GPU, the frame loop
 * that the reference distributes over MPI ranks (animation.cpp:44-100 with
 * uniform_work_partitioner.cpp:36-52; illcrawl_render.cpp:189-229) and its
 * per-frame projection::create_projection (projection.cpp:49-113). */

typedef struct ilc_batch ilc_batch;

typedef enum {
  ILC_QUANTITY_MASS = 0, /* w = particle mass: column density
                            (unprocessed_quantity, quantities.cl:701-711) */
  ILC_QUANTITY_XRAY = 1  /* w = bolometric free-free emissivity (see
                            ilc_quantity_xray_bolometric) */
} ilc_quantity_kind;

typedef struct {
  ilc_camera cam;
  int32_t quantity; /* ilc_quantity_kind */
  float scale;      /* effective dA applied at retrieval
                       (projective_smoothing_reconstruction.cpp:49-84) */
} ilc_view;

/* ilc_batch_render flags */
#define ILC_RENDER_PLAIN 1 /* one gather launch per view (the plain path) */
#define ILC_RENDER_GRAPH 2 /* replay the device work from a hipGraph while the
                              plan (every view's grid) is unchanged */

ilc_status ilc_batch_create(const ilc_psg_config* cfg, ilc_batch** out);
ilc_status ilc_batch_destroy(ilc_batch* b);
/* Resident snapshot, borrowed until the next call: pos3_dev (n x, y, z
 * interleaved), mass_dev, hsml_dev; rho/temp/xe (may be NULL unless an
 * X-ray view is rendered).  The crawler's per-block particle setup
 * (reconstructing_data_crawler.cpp:164-255) with the whole snapshot as one
 * block. */
ilc_status ilc_batch_set_particles(ilc_batch* b, const float* pos3_dev, const float* mass_dev,
                                   const float* hsml_dev, const float* rho_dev,
                                   const float* temp_dev, const float* xe_dev, uint64_t n);
/* 1-D Gaunt table for X-ray views (node i at x0 + i*dx on ln T). */
ilc_status ilc_batch_set_gaunt_table(ilc_batch* b, const float* table_host, int32_t ntab,
                                     float x0, float dx);
/* Views; computes each view's pixel points (camera.cpp:157-189). */
ilc_status ilc_batch_set_views(ilc_batch* b, const ilc_view* views, int32_t nviews,
                               void* stream);
/* Render every view (synchronises the stream before returning). */
ilc_status ilc_batch_render(ilc_batch* b, int32_t flags, void* stream);
/* Image of a view: nx*ny floats (ix + nx*iy), scaled by the view's dA. */
ilc_status ilc_batch_image(const ilc_batch* b, int32_t view, float* out_dev, void* stream);
/* Its FP64 accumulator (unscaled). */
ilc_status ilc_batch_image_f64(const ilc_batch* b, int32_t view, double* out_dev, void* stream);
/* Grid of a view after a render (the particle_grid getters,
 * particle_grid.cpp:121-134, :342-346).  num_particles counts the particles
 * the view's grid holds: a view whose screen lies well inside the particle
 * bounding box keeps only the particles of the cells its points can reach
 * (the gather never visits the others, so its image is unchanged). */
ilc_status ilc_batch_view_info(const ilc_batch* b, int32_t view, ilc_grid_info* out);
/* Exact particle-kernel evaluation count of each view (after a render). */
ilc_status ilc_batch_count_evals(ilc_batch* b, uint64_t* per_view_host, void* stream);
/* Adjoint of the last render (exact mode; the reference has no counterpart).
 * dl_dimage_dev[v]: dL/d(image of view v) (nx*ny floats, device, as
 * ilc_batch_image returns it, i.e. with the view's dA scale).  Writes per
 * snapshot particle (n entries, device, overwritten):
 *   dl_dw:    sum over views of dL/dw_v, w_v the particle's weight in view v
 *             (its mass, or its X-ray emissivity), 0 from views whose depth
 *             cut removes the particle;
 *   dl_dh:    dL/dh summed over views;
 *   dl_dpos3: dL/d(x, y, z) in world coordinates (3 n floats, xyz interleaved).
 * Gradients are FP32 atomic sums (order varies from call to call).
 * Synchronises the stream. */
ilc_status ilc_batch_backward(ilc_batch* b, const float* const* dl_dimage_dev, float* dl_dw_dev,
                              float* dl_dh_dev, float* dl_dpos3_dev, void* stream);
/* Device time (ms, HIP events on the render stream) of the gather launch(es)
 * of the last non-graph render: the one persistent launch, or the sum of the
 * per-view launches of a plain render.  Measurement only. */
ilc_status ilc_batch_last_gather_ms(const ilc_batch* b, float* ms);
/* Gather tile (points); tile_w * tile_h must be a multiple of 256.  One
 * tile per workgroup of 256 threads. */
ilc_status ilc_batch_set_tile(ilc_batch* b, int32_t tile_w, int32_t tile_h);
/* The same with the workgroup size: 256 threads (the shapes of
 * ilc_batch_set_tile), or 64 (one wave per 16x16 tile, four points per
 * lane: four times the tiles in flight per CU, for latency-bound sparse
 * views).  No reference counterpart: tuning of this build's gather. */
ilc_status ilc_batch_set_tile_wg(ilc_batch* b, int32_t tile_w, int32_t tile_h,
                                 int32_t workgroup);
/* The current tile shape and workgroup size (after ilc_batch_autotune). */
ilc_status ilc_batch_get_tile(const ilc_batch* b, int32_t* tile_w, int32_t* tile_h,
                              int32_t* workgroup);
/* Depth cut of every view.  clip = 0 (default): the reference's centre
 * mask, a particle counts whole when its camera depth lies in [0, depth] and
 * not at all otherwise (projection.cl:80-91).  clip = 1: the line-of-sight
 * kernel clip, which the reference does not have: each particle adds the
 * part of its 3-D quartic kernel's column that lies in [0, depth] (slabs of
 * a tomography stack share the particles their supports straddle).  Needs
 * exact mode and a quartic kernel; renders must be persistent (not plain),
 * and backward is not provided for it. */
ilc_status ilc_batch_set_depth_clip(ilc_batch* b, int32_t clip);
/* Times the batched render for each supported tile shape, keeps the
 * fastest and returns it (no reference counterpart: measurement-driven
 * tuning). */
ilc_status ilc_batch_autotune(ilc_batch* b, int32_t* tile_w, int32_t* tile_h, double* best_ms,
                              void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ILLCRAWL_PSG_H */
