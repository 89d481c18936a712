/*
 * tr_api.h — C ABI of the MI355X transient path tracer.
 *
 * The library replaces the reference renderer's transient path-tracing hot path
 * (TransientRenderer::render -> World::Li -> BidirectionalPathTracing in
 * path-tracing-only mode -> StreakCameraFilm::add_samples) with a wavefront
 * path tracer of hand-written HIP kernels for gfx950.
 *
 * Conventions
 *   - Every entry point returns TR_OK (0) or a negative TR_ERR_* status; no
 *     exceptions cross the ABI. tr_last_error() returns a message for the last
 *     failure on the calling thread.
 *   - Host arrays passed in are caller-owned and copied. Device memory is owned
 *     by the scene / film handle that allocated it.
 *   - A handle is used by one host thread at a time; distinct handles may be
 *     used concurrently. Work is issued on the library's own HIP stream per
 *     handle and every entry point returns after that work has completed,
 *     except where noted.
 *   - The library never falls back to CPU code: without a usable gfx950 device,
 *     tr_scene_create / tr_film_create fail with TR_ERR_DEVICE.
 *
 * Reference interfaces each entry point replaces are cited as file:line in the
 * reference repository (src/...).
 */
#ifndef TR_API_H
#define TR_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TR_API_VERSION 1

enum {
    TR_OK = 0,
    TR_ERR_INVALID = -1,   /* bad argument / handle state */
    TR_ERR_DEVICE = -2,    /* HIP runtime or kernel failure, or no gfx950 device */
    TR_ERR_NOMEM = -3,     /* device or host allocation failed */
    TR_ERR_STATE = -4      /* call order violated (e.g. render before commit) */
};

typedef struct tr_scene_s* tr_scene;
typedef struct tr_film_s* tr_film;

/* Lambertian material (UberMaterial + one Lambertian BSDF: bunnykiller.cpp:328-336,
 * src/Material/UberMaterial.h:71-104, src/Material/Reflectance/Lambertian.h:60-122). */
typedef struct {
    float kd[3];
} tr_material;

/* ---- scene (World) ----------------------------------------------------------------------- */

/* World::World (src/RayTracing/World.h:137-157). */
int tr_scene_create(tr_scene* out);
int tr_scene_destroy(tr_scene scene);

/* World::add_triangle_mesh (src/RayTracing/World.h:209-541). Triangles keep the reference's
 * conventions: geometric/shading normal normalize(cross(p2-p0, p1-p0)) (World.h:449);
 * geom_id = order of mesh insertion, prim_id = triangle order within the mesh.
 * xyz: nv*3 floats, idx: nt*3 vertex indices. */
int tr_scene_add_mesh(tr_scene scene, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                      const tr_material* material, uint32_t* geom_id);

enum {
    /* -two-sided: TwoSidedMaterial around the Lambertian (src/Material/TwoSided.h:34-96): hit from
     * behind, the shading frame is inverted, and cosines are taken as |cos| (Path.h:188-193) */
    TR_MATERIAL_TWO_SIDED = 1u
};

/* tr_scene_add_mesh with per-corner shading normals (`normals`: 9 floats per triangle, corners in
 * index order; NULL = the geometric normal) interpolated at each hit as the reference does with
 * OBJ "vn" data (World.h:490-509, 589-598; not renormalised), and material flags. */
int tr_scene_add_mesh_ex(tr_scene scene, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                         const float* normals, const tr_material* material, uint32_t flags, uint32_t* geom_id);

/* -name-mesh file.obj: World::add_triangle_mesh (src/RayTracing/World.h:209-541) reading a
 * Wavefront OBJ (triangles only; "f v", "v/vt", "v//vn", "v/vt/vn" corners; "vn" normals are
 * normalised and used as shading normals). Errors carry the reference's messages. */
int tr_scene_add_obj(tr_scene scene, const char* path, const tr_material* material, uint32_t flags,
                     uint32_t* geom_id);

/* The OBJ reader alone (host only): vertices, triangle indices and, if the file has "vn" data,
 * per-corner normals (9 floats per triangle; *normals = NULL otherwise). Release with tr_obj_free. */
int tr_obj_load(const char* path, float** xyz, uint32_t* nv, uint32_t** idx, uint32_t* nt, float** normals);
int tr_obj_free(float* xyz, uint32_t* idx, float* normals);

/* PointLightSource (src/LightSource/PointLightSource.h:39-171; CLI bunnykiller.cpp:1194-1207). */
int tr_scene_add_point_light(tr_scene scene, const float position[3], const float intensity[3]);

/* RectangularAreaLightSource with the CLI parameterisation (bunnykiller.cpp:1242-1261):
 * corner, a target point (normal = normalize(target - corner)), side lengths, scalar intensity.
 * (src/LightSource/RectangularAreaLightSource.h:66-80, 222-248). */
int tr_scene_add_rect_light(tr_scene scene, const float corner[3], const float target[3],
                            const float size[2], float intensity);

/* World::set_medium with HomogeneousMedium<Isotropic> (src/RayTracing/World.h:711-715,
 * src/Media/HomogeneousMedium.h:45-127, bunnykiller.cpp:178-228). No time delay. */
int tr_scene_set_medium(tr_scene scene, const float sigma_a[3], const float sigma_s[3]);

/* World::freeze (src/RayTracing/World.h:543-550): builds the BVH and uploads the scene. */
int tr_scene_commit(tr_scene scene);

typedef struct {
    uint32_t n_tris, n_bvh_nodes, bvh_depth, n_lights;
    uint64_t device_bytes;
    double build_ms;
} tr_scene_info;
int tr_scene_info_get(tr_scene scene, tr_scene_info* info);

/* ---- ray queries (World::first_intersection / intersects) -------------------------------- */

typedef struct {
    float org[3];
    float tnear;
    float dir[3];
    float tfar;
} tr_ray; /* 32 bytes */

typedef struct {
    float t;          /* hit distance, or -1 on a miss */
    uint32_t geom_id; /* 0xFFFFFFFF on a miss (RTC_INVALID_GEOMETRY_ID) */
    uint32_t prim_id;
    uint32_t tri;     /* scene-wide triangle index */
} tr_hit; /* 16 bytes */

/* Closest hit in (tnear, tfar] for each ray — World::first_intersection's query
 * (src/RayTracing/World.h:566-615 -> rtcIntersect). Watertight ray/triangle test.
 * Host pointers; n may be 0. */
int tr_intersect_batch(tr_scene scene, const tr_ray* rays, uint32_t n, tr_hit* hits);
/* Same, on device pointers, asynchronously on `stream` (a hipStream_t; NULL = library stream). */
int tr_intersect_batch_device(tr_scene scene, const tr_ray* d_rays, uint32_t n, tr_hit* d_hits,
                              void* stream);
/* tr_intersect_batch without allocating device memory. No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). Rays and hits are staged in a static device array of TR_NOALLOC_MAX_RAYS x 48 bytes
 * (3 MiB) on each GPU the library runs on, held until the process ends; n above
 * TR_NOALLOC_MAX_RAYS fails with TR_ERR_INVALID. Every call on one GPU shares that array: calls
 * that overlap on one GPU (from several threads) race and must be serialised by the caller.
 * Host pointers; same results as tr_intersect_batch. */
#define TR_NOALLOC_MAX_RAYS 65536u
int tr_intersect_batch_noalloc(tr_scene scene, const tr_ray* rays, uint32_t n, tr_hit* hits);

/* Any hit in (tnear, tfar] for each ray: occluded[i] = 1 if some triangle is hit there, else 0 —
 * World::intersects (src/RayTracing/World.h:617-639 -> rtcIntersect, geomID test) and
 * World::is_visible (World.h:641-664 -> rtcOccluded; visible = !occluded for the segment
 * v1 -> v2 with tnear = epsilon, tfar = |v2 - v1| - epsilon). Same watertight test as
 * tr_intersect_batch, so occluded[i] == (tr_intersect_batch finds a hit) for every ray.
 * Host pointers; n may be 0. */
int tr_occluded_batch(tr_scene scene, const tr_ray* rays, uint32_t n, uint8_t* occluded);
/* Same, on device pointers, asynchronously on `stream` (a hipStream_t; NULL = library stream). */
int tr_occluded_batch_device(tr_scene scene, const tr_ray* d_rays, uint32_t n, uint8_t* d_occluded,
                             void* stream);

/* ---- film (StreakCameraFilm) ------------------------------------------------------------- */

typedef struct {
    int32_t width, height, time_bins;
    float exposure;       /* -film-exposure */
    float offset;         /* -film-offset (StreakCameraFilm::set_offset) */
    int32_t scanline;     /* -film-scanline y, or -1 for the whole film */
    int32_t camera_unwarp;/* -camera-unwarp */
    /* -multibounce-streak N: MultibounceStreakFilm (src/Film/MultibounceStreakCameraFilm.h:67-171)
     * with N bounce films (a sample of bounce b >= N goes to the last one); 0 = one streak film */
    int32_t bounce_layers;
} tr_film_desc;

/* StreakCameraFilm::StreakCameraFilm (src/Film/StreakCameraFilm.h:91-121). Device buffers: a
 * compensated FP32 sum, two arrays of rows x W x T x 3 floats (rows = 1 in scanline mode): the
 * running sum and the exact rounding errors of the additions into it. Every addition into an
 * element (each sample of a render, each FP64 LDS partial the film kernel flushes, each caller
 * sample) is compensated, so sum + comp is within (2u + O(n u^2)) sum|x| of the exact total after n
 * additions (u = 2^-24), against n u sum|x| for a plain FP32 running sum, whatever the order of the
 * additions. A staging array for the normalised copy is allocated by the first read. */
int tr_film_create(tr_film* out, const tr_film_desc* desc);
int tr_film_destroy(tr_film film);
/* Film::clean: zero all sums and weights. */
int tr_film_clear(tr_film film);
/* Normalised streak data, StreakCameraFilm::average_and_save_slice (StreakCameraFilm.h:357-397)
 * with every pixel divided by its own weight (0/0 -> 0): out[row][x][t][c], rows x W x T x 3.
 * steady (optional, may be NULL): the time-integrated image out[row][x][c] (Film::draw_pixel).
 * Multibounce films: out[layer][row][x][t][c], each bounce film divided by its own weights (the
 * camera samples that sent it samples, MultibounceStreakCameraFilm.h:137-157); steady holds every
 * sample, inside the time window or not (Film::add_samples, Film.h:263-283). */
int tr_film_read(tr_film film, float* out, float* steady);
/* tr_film_read for the rows [row_begin, row_begin + row_count) of the film only (row index 0 is
 * the scanline in scanline mode): out[row][x][t][c] and steady[row][x][c] of those rows, as
 * StreakCameraFilm::average_and_save_slice produces them one scanline at a time
 * (StreakCameraFilm.h:357-397). Not for multibounce films. */
int tr_film_read_rows(tr_film film, int32_t row_begin, int32_t row_count, float* out, float* steady);
/* Un-normalised compensated sums (sum + compensation) as double, and per-pixel weights
 * ([layer][row][x] for multibounce films). */
int tr_film_read_sums(tr_film film, double* sums, double* weights);
/* Device pointers of the compensated-sum buffers (read-only use by the caller). */
int tr_film_device_buffers(tr_film film, float** d_sum, float** d_comp);

/* One time-resolved radiance sample for tr_film_add_samples (RadianceSample, RadianceSample.h:34-73,
 * with the film pixel its camera sample lies in). */
typedef struct {
    int32_t x, y;     /* film pixel; y is the image row (the scanline itself in scanline mode) */
    float time;       /* path length: the sample's time, already unwarped if the caller wants that */
    float rgb[3];     /* radiance, already weighted by the camera sample's filter weight */
    float weight;     /* added once to the pixel's weight (Film::m_weight): the camera sample's
                       * weight on one of its samples, 0 on the others */
    uint32_t pad0;
} tr_film_sample; /* 32 bytes */

/* StreakCameraFilm::add_samples (src/Film/StreakCameraFilm.h:139-192, 280-297) for samples the
 * caller made: bin = floor((time - offset) / exposure), samples outside [0, T) are not deposited
 * (their weight still counts), each deposit is a compensated addition by the film kernel that
 * renders use. Samples outside the film are an error (nothing is added). Not for multibounce
 * films. Host pointer; n <= 2^31. */
int tr_film_add_samples(tr_film film, const tr_film_sample* samples, uint32_t n);
/* StreakCameraFilm::write (src/Film/StreakCameraFilm.h:233-242, 357-437) and Film::write
 * (src/Film/Film.h:358-471) with Imaging::save (src/Image/ImageIO.h:147-194): one streak image
 * per scanline y, "<name>_%04d.<ext>" (T wide, W high; stored transposed when T < 8), ext "hdr"
 * (Radiance RGBE) or "raw" (FP32 dump), then the time-integrated steady image "<name>.hdr". */
int tr_film_write(tr_film film, const char* name, const char* ext);
/* Imaging::save (src/Image/ImageIO.h:147-194) for an RGB float image [height][width][3]: the
 * extension of `path` selects Radiance RGBE (.hdr) or a raw FP32 dump (.raw). Host only. */
int tr_image_save(const char* path, const float* rgb, int width, int height);

/* ---- rendering (TransientRenderer::render) ------------------------------------------------ */

typedef struct {
    float position[3];
    float focus[3];
    float up[3];
    float fov_deg;         /* -camera-fov: view-plane distance 1/tan(fov/2) (bunnykiller.cpp:1156-1160) */
    float view_plane_dist; /* PinholeCamera's fd (PinholeCamera.h:63-80); used instead of fov_deg if > 0 */
} tr_camera;

typedef struct {
    int32_t sqrt_spp;     /* -camera-spp: sqrt of camera samples per pixel (StratifiedSampler) */
    int32_t n_incoming;   /* -bidirectional-path-tracing N: paths per camera sample */
    int32_t max_bounces;  /* -max-nb-bounces */
    int32_t pad0;
    uint64_t seed;        /* -new-seed */
    uint32_t wave_paths;  /* paths in flight per wavefront (0 = library default) */
    uint32_t flags;       /* TR_RENDER_* */
} tr_render_desc;

enum {
    TR_RENDER_PROFILE = 1u,     /* time each kernel with HIP events (tr_render_stats.kernel_ms) */
    TR_RENDER_BINARY_BVH = 2u,  /* traverse the binary BVH instead of the default 4-wide one */
    /* sample Lambertian directions in an orthonormal frame around the normal instead of the
     * reference's frame (Intersection.h:167-178 builds a tangent that is not orthogonal to n
     * when n.z != 0); guided renders always use the orthonormal frame */
    TR_RENDER_ORTHONORMAL_FRAME = 4u,
    /* capture the whole render — every wavefront's camera and bounce kernels — in one hipGraph
     * and submit it with a single launch (guided renders: one wavefront's kernels, replayed per
     * wavefront around the training steps); per-kernel timing is not available in this mode */
    TR_RENDER_GRAPH = 8u,
    /* the integrator's Transient sampling mode, -time-sampling (BidirectionalPathTracing.h:190-198,
     * bunnykiller.cpp:1811-1817): free flights at rate 4 / (T * exposure), directions at medium
     * vertices drawn by time-angular sampling towards a light point chosen per path
     * (SamplingTechniques.h:170-250), no Russian roulette (paths end at the film's time length or
     * the bounce cap). A per-job flag: jobs of one batch may differ. Not combined with guiding. */
    TR_RENDER_TIME_SAMPLING = 16u,
    /* traversal stack entries of the 4-wide BVH (launch tuning; results do not change). By default
     * a BVH of at most 65,536 nodes uses 16-bit node indices (half the LDS of 32-bit ones, so more
     * wavefronts per compute unit). STACK32: 32-bit entries always. STACK_KEYED: closest-hit
     * stacks keep each entry's box distance beside its 16-bit index, and entries behind the
     * current hit are dropped without visiting them. */
    TR_RENDER_STACK32 = 32u,
    TR_RENDER_STACK_KEYED = 64u,
    /* TR_RENDER_PROFILE for the closest-hit (extend) launches only: kernel_ms of the other kinds
     * stays 0. Each event pair between two launches drains the GPU, so timing every kernel adds
     * 2-14 % to a render; timing one kind adds a fraction of that. */
    TR_RENDER_PROFILE_EXTEND = 256u,
    /* batches: render each job of at least 2^22 paths on its own (its own wavefronts, single-job
     * kernels and the LDS-privatised film), one after another, and the smaller jobs together in one
     * shared sequence of wavefronts; without the flag every job shares one sequence. tr_autotune
     * tries both */
    TR_RENDER_PER_JOB = 128u,
    /* closest-hit launches run node steps and triangle steps in separate wavefront-wide phases: a
     * node's entered leaves wait until many lanes hold some (keyed 16-bit stack; BVHs of at most
     * 65,536 nodes, otherwise ignored). Launch tuning; results do not change. */
    TR_RENDER_POSTPONE_LEAVES = 512u,
    /* camera rays (the first bounce of a single-job render) traced one query per lane like the
     * later bounces, instead of as packets of one pixel's 64 samples per wavefront. Launch tuning
     * and testing; results do not change. */
    TR_RENDER_NO_CAMERA_PACKETS = 1024u
};

enum {
    /* TR_K_FILM_MERGE: no longer launched (films accumulate compensated sums directly); kept so the
     * statistics layout does not change */
    TR_K_CAMERA = 0, TR_K_EXTEND = 1, TR_K_SHADE = 2, TR_K_SHADOW = 3, TR_K_FILM_MERGE = 4, TR_K_FILM = 5,
    TR_K_GUIDE = 6, /* guided renders: training-record bookkeeping and Adam steps */
    TR_K_COUNT = 8
};

typedef struct {
    uint64_t paths, extension_rays, shadow_rays, contributions;
    uint32_t waves, bounces_max;
    double render_ms;              /* host wall time of the whole call */
    double kernel_ms[TR_K_COUNT];  /* summed HIP-event time per kernel kind (TR_RENDER_PROFILE) */
    uint32_t kernel_launches[TR_K_COUNT];
} tr_render_stats;

/* TransientRenderer::render (src/Renderer/TransientRenderer.h:77-201) driving
 * BidirectionalPathTracing::operator()(ray, samples, ...) in path-tracing-only SteadyState mode
 * (src/Integrator/BidirectionalPathTracing.h:1100-1140) and StreakCameraFilm::add_samples
 * (src/Film/StreakCameraFilm.h:139-192). Accumulates into `film` with compensated additions
 * (successive calls add further passes). Non-finite samples are zeroed (TransientRenderer.h:167-171).
 * Random numbers: one PCG32 stream per camera sample, per path and per path for light sampling,
 * all seeded from desc->seed. */
int tr_render_transient(tr_scene scene, const tr_camera* camera, tr_film film,
                        const tr_render_desc* desc, tr_render_stats* stats);

/* ---- batched renders (no counterpart in the reference) -------------------------------------
 * Many render jobs (scene, camera, film, sampling), typically of uneven size, rendered by ONE
 * sequence of wavefronts: paths of all jobs share the persistent traversal and shading kernels,
 * and each path looks up its job, so small jobs do not leave the GPU idle. Scenes and films are
 * referenced, not copied: keep them alive until tr_batch_render returns. Each job accumulates
 * into its own film as tr_render_transient would (films must be distinct). */
typedef struct tr_batch_s* tr_batch;

typedef struct {
    uint32_t wave_paths;    /* paths per wavefront (0 = library default) */
    uint32_t flags;         /* TR_RENDER_PROFILE | TR_RENDER_BINARY_BVH | TR_RENDER_ORTHONORMAL_FRAME | TR_RENDER_GRAPH
                             * | TR_RENDER_STACK32 | TR_RENDER_STACK_KEYED | TR_RENDER_PER_JOB */
    uint32_t trace_per_cu;  /* persistent traversal blocks per compute unit (0 = default 24) */
    uint32_t shade_per_cu;  /* shading blocks per compute unit (0 = default 2) */
} tr_launch_desc;

int tr_batch_create(tr_batch* out);
int tr_batch_destroy(tr_batch batch);
/* Appends a job; `job` (optional) receives its index. Jobs render in the order added. */
int tr_batch_add(tr_batch batch, tr_scene scene, const tr_camera* camera, tr_film film,
                 const tr_render_desc* desc, uint32_t* job);
int tr_batch_clear(tr_batch batch);
/* Renders every job; `launch` may be null (defaults, or the last tr_autotune result).
 * stats aggregates all jobs. */
int tr_batch_render(tr_batch batch, const tr_launch_desc* launch, tr_render_stats* stats);

/* Autotuner: times candidate launch shapes on the batch's own jobs and keeps the fastest for later
 * tr_batch_render calls (returned in `best`, optional; its trial time in `ms_best`, optional).
 * Trials render every job with sqrt_spp capped at the smallest value for which the capped batch
 * holds min(all paths, 2^25) paths (about one largest-candidate wavefront), best of two runs each:
 * stage 1 wavefront sizes 2^18..2^25 with and without graph replay, stage 2 traversal / shading
 * blocks per CU, stage 4 the traversal formats (32-bit stack, keyed stack, postponed leaves).
 * Stage 3 renders the FULL jobs once in each batch mode (shared wavefronts; TR_RENDER_PER_JOB), so
 * the call costs at least two complete batch renders plus the trials. A candidate replaces the
 * incumbent only when more than 2 % faster. Every film is cleared afterwards. */
int tr_autotune(tr_batch batch, tr_launch_desc* best, double* ms_best);

/* ---- path guiding (no counterpart in the reference) ---------------------------------------
 * A 48 -> 64 -> 64 -> 64 -> 64 -> 40 ReLU network (BF16 weights, FP32 accumulation on the matrix
 * cores) maps a path vertex (position, incoming direction, normal, time, medium flag) to a mixture
 * of 8 von Mises-Fisher lobes. Guided renders draw each next direction from the one-sample MIS
 * mixture bsdf_fraction * BSDF/phase + (1 - bsdf_fraction) * network, so the film stays an unbiased
 * estimate, and train the network online (Adam) on the radiance found along sampled directions.
 * A guide keeps its weights across renders. */
typedef struct tr_guide_s* tr_guide;

typedef struct {
    float learning_rate;     /* Adam step size (default 1e-3) */
    float bsdf_fraction;     /* probability of sampling the BSDF / phase function, in (0, 1] (0.5) */
    uint32_t train_paths;    /* paths per wavefront whose vertices become training records (65536) */
    uint32_t batch_size;     /* records per Adam step, a multiple of 64, at most 2^20 (16384) */
    uint32_t steps_per_wave; /* Adam steps after each wavefront (1) */
    uint32_t pad0;
    uint64_t init_seed;      /* deterministic weight initialisation (1) */
} tr_guide_desc;

typedef struct {
    uint64_t steps;          /* Adam steps taken */
    uint64_t records;        /* records passed to tr_guide_train_batch */
    double last_loss;        /* mean loss of the last step */
    uint32_t n_params;       /* FP32 master parameters (18,216) */
    uint32_t flop_per_eval;  /* network FLOP per evaluated vertex (35,840) */
} tr_guide_info;

/* Creates a guide (desc may be null for the defaults). */
int tr_guide_create(tr_guide* out, const tr_guide_desc* desc);
int tr_guide_destroy(tr_guide guide);
/* Master FP32 parameters, per layer W[in][out] (in-major) then b[out]; setting them resets Adam. */
int tr_guide_get_params(tr_guide guide, float* params);
int tr_guide_set_params(tr_guide guide, const float* params);
/* Adam steps after each wavefront of later guided renders; 0 freezes the network (renders then
 * collect no training records). */
int tr_guide_set_steps(tr_guide guide, uint32_t steps_per_wave);
/* Network outputs for n raw vertex inputs (12 floats each: position normalised to the scene box,
 * incoming direction wo, normal (zero in a medium), time / film length, medium flag, 0); out holds
 * n x 40 floats (lobe i: direction xyz, concentration and weight logits at 5i..5i+4). */
int tr_guide_eval(tr_guide guide, const float* raw, uint32_t n, float* out);
/* Same on device pointers; kernel_ms (optional) receives the kernel time from HIP events. */
int tr_guide_eval_device(tr_guide guide, const float* d_raw, uint32_t n, float* d_out, float* kernel_ms);
/* Mean loss gradient over n <= batch_size training records (16 floats each: position xyz, t,
 * wo xyz, medium, normal xyz, weight, sampled direction xyz, 0); weights are normalised by their
 * mean as in training. grad receives n_params floats. */
int tr_guide_gradient(tr_guide guide, const float* records, uint32_t n, float* grad, float* loss);
/* One Adam step on n <= batch_size records. */
int tr_guide_train_batch(tr_guide guide, const float* records, uint32_t n, float* loss);
int tr_guide_info_get(tr_guide guide, tr_guide_info* info);

/* tr_render_transient with guided direction sampling (orthonormal shading frame) and online
 * training of `guide` after every wavefront. */
int tr_render_transient_guided(tr_scene scene, const tr_camera* camera, tr_film film,
                               const tr_render_desc* desc, tr_guide guide, tr_render_stats* stats);

/* ---- misc --------------------------------------------------------------------------------- */

const char* tr_last_error(void);
int tr_device_info(char* name, int name_len, int* compute_units, uint64_t* global_mem);

#ifdef __cplusplus
}
#endif
#endif /* TR_API_H */
