/*
 * wave-tool water-displacement stage for MI355X (gfx950) — C ABI.
 *
 * The reference evaluates this stage in an OpenGL vertex shader
 * (assets/shaders/water-grid.vert) driven by the water block of
 * RenderEngine::render (src/render-engine.cpp:796-1123). That block's host half
 * (projector and uniforms) and device half (per-vertex Gerstner sum, heightmap
 * bump, normals) are replaced by the entry points below. Each declaration names
 * the reference interface it replaces (paths relative to the reference root).
 *
 * Conventions
 *  - Plain C types only; device pointers are plain pointers, streams are
 *    `void*` holding a hipStream_t (NULL = the default stream).
 *  - Every entry point returns a wt_status; no exception crosses the ABI.
 *    wt_last_error() returns the message of the last failing call on a context.
 *  - Calls on one context are stream-ordered; one context per host thread.
 *  - Vertex output layout is the reference's MeshObject attribute layout
 *    (src/mesh-object.h:74-75, src/render-engine.cpp:1175-1191): tightly packed
 *    float3 arrays (12-byte stride), vertex id = row * gridLength + column with
 *    row 0 the bottom edge of the grid (water-grid.vert:115-125).
 */
#ifndef WT_H_
#define WT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WT_ABI_VERSION 1

/* Largest wave count one call accepts (the reference fixes 4,
 * water-grid.vert:35 and src/render-engine.h:132). */
#define WT_MAX_WAVES 1024

/* Algorithmic work units used for roofline accounting (not executed counts).
 * Forward: 26 FP32 FLOP per vertex-wave (FMA = 2, sin and cos = 1 each):
 * phase 5, sin+cos 2, position 7, analytic partials 12. Bytes: HBM writes of
 * position + normal (24 B per vertex) or with tangent + bitangent (48 B).
 * Adjoint: 50 FLOP per vertex-wave-snapshot per iteration. */
#define WT_FLOP_PER_VERTEX_WAVE 26
#define WT_BYTES_PER_VERTEX_PN 24
#define WT_BYTES_PER_VERTEX_PNTB 48
#define WT_ADJOINT_FLOP_PER_VERTEX_WAVE_SNAPSHOT 50

typedef enum wt_status {
  WT_OK = 0,
  WT_ERR_INVALID_ARG = 1,
  WT_ERR_NO_INTERSECTION = 2, /* frustum misses the displaceable slab (render-engine.cpp:894) */
  WT_ERR_HIP = 3,
  WT_ERR_OOM = 4,
  WT_ERR_UNSUPPORTED = 5,
  WT_ERR_DEGENERATE = 6 /* a projector ray parallel to y = 0 (render-engine.cpp:934-938, 1034 asserts) */
} wt_status;

typedef struct wt_ctx wt_ctx;
typedef struct wt_batch wt_batch;
typedef struct wt_graph wt_graph;
typedef struct wt_fit wt_fit;

/* One Gerstner wave: the GLSL uniform struct GerstnerWave
 * (water-grid.vert:45-51) in its member order, plus an appended phase offset
 * phi0 added to the phase (zero reproduces the reference). steepness_Q_i is
 * the already-normalised Q_i the shader receives (see wt_normalize_steepness).
 * Passed to the kernels as a device array. */
typedef struct wt_gerstner_wave {
  float amplitude_A;
  float frequency_w;
  float phaseConstant_phi;
  float steepness_Q_i;
  float xzDirection_D[2];
  float phaseOffset_phi0;
  float reserved;
} wt_gerstner_wave;

/* Camera state of src/camera.h:60-72 that the projector reads. The camera is
 * built with the reference defaults (forward (0,0,-1), up (0,1,0),
 * src/camera.cpp:40-45) and then setRotation(yaw_deg, pitch_deg) is applied. */
typedef struct wt_camera {
  float fov_deg;
  float aspect;
  float near_clip;
  float far_clip;
  float position[3];
  float yaw_deg;
  float pitch_deg;
} wt_camera;

typedef enum wt_wrap { WT_WRAP_REPEAT = 0, WT_WRAP_MIRRORED_REPEAT = 1 } wt_wrap;
typedef enum wt_texel_format { WT_TEXEL_R8_UNORM = 0, WT_TEXEL_R32F = 1 } wt_texel_format;

/* Heightmap sampler state (src/texture.cpp:54-72): bilinear GL_LINEAR
 * filtering with texel centres at (i + 1/2) / size, wrap REPEAT or
 * MIRRORED_REPEAT (the reference uses MIRRORED_REPEAT, texture.cpp:63-64).
 * texels: device pointer, row-major, row 0 = texture t in [0, 1/height)
 * (GL memory order, i.e. after the loader's vertical flip,
 * render-engine.cpp:1310). R8_UNORM holds the .r channel the shader reads
 * (water-grid.vert:107) as one byte per texel. */
typedef struct wt_heightmap {
  const void* texels;
  uint32_t width;
  uint32_t height;
  uint32_t format; /* wt_texel_format */
  uint32_t wrap;   /* wt_wrap */
} wt_heightmap;

/* The water program's uniform set (water-grid.vert:37-64, uploaded at
 * render-engine.cpp:1057-1110). Corners are world positions on y = 0 in the
 * order of waterGridCornerPoints (render-engine.cpp:1012-1015): [0] bottom-left,
 * [1] top-left, [2] bottom-right, [3] top-right, each (x, y, z, w); they are
 * double so the FP64 phase reduction sees exact grid positions. The scroll
 * velocity moves the heightmap sample point by -v*t (zero in the reference). */
typedef struct wt_water_params {
  double corners[4][4];
  float cameraPosition[3];
  uint32_t gridLength;
  float heightmapSampleScale;
  float heightmapDisplacementScale;
  float verticalBounceWaveDisplacement;
  float heightmapScrollVelocity[2];
  float reserved;
  double waveAnimationTimeInSeconds;
} wt_water_params;

/* Caller-owned device output arrays, each gridLength^2 float3 (12-byte stride).
 * position and normal are required; tangent and bitangent may be NULL. */
typedef struct wt_vertex_buffers {
  float* position;
  float* normal;
  float* tangent;
  float* bitangent;
} wt_vertex_buffers;

/* Kernel launch shape for the forward kernels; 0 in a field means "autotuned
 * or default". */
typedef struct wt_launch_config {
  uint32_t wave_split;     /* wavefronts sharing one block of vertices, each summing 1/wave_split of the waves: 1 or 2 */
  uint32_t steps_per_lane; /* vertices each lane advances along its row by the sin/cos recurrence: 4, 8 or 16 */
  uint32_t lanes_per_row;  /* lanes of a 64-wide wavefront sharing one grid row: 64, 32, 16 or 8 */
  uint32_t wave_chunk;     /* waves staged in LDS per pass: 64, or 32 / 128 with wave_split 2 and 64 lanes per row */
} wt_launch_config;

/* ---------------------------------------------------------------- context */
int wt_abi_version(void);
const char* wt_status_string(wt_status s);
/* Binds to the HIP device current on the calling thread. */
wt_status wt_create(wt_ctx** out);
void wt_destroy(wt_ctx* ctx);
const char* wt_last_error(const wt_ctx* ctx);

/* ---------------------------------------------------------------- host helpers */
/* Camera::updateBasisVectors / updateProjectionMat / updateViewMat
 * (src/camera.cpp:130-146): column-major view and projection matrices and the
 * forward vector, in double. */
wt_status wt_camera_matrices(const wt_camera* cam, double view[16], double proj[16], double forward[3]);

/* Projector: the host half of the water block, src/render-engine.cpp:797-1042
 * (with Line/Plane/linePlaneIntersection, src/render-engine.h:53-124, and the
 * Camera clone). Returns the four world-space grid corners in the
 * wt_water_params.corners order. displaceable_amplitude is the slab half-height
 * of render-engine.cpp:803 (passed explicitly; see the reference's stale
 * TotalAmplitude, render-engine.h:156-162). Computed in double.
 * WT_ERR_NO_INTERSECTION where the reference silently skips the water. */
wt_status wt_projector_corners(const wt_camera* cam, float displaceable_amplitude, double corners[4][4]);

/* The same projector (src/render-engine.cpp:797-1042) on the device for many
 * cameras, one thread per camera, in double. d_cameras: count wt_camera structs;
 * d_amplitudes: count slab half-heights. Writes d_corners[count][4][4] (NaN for a
 * camera without corners) and d_status[count] (WT_OK, WT_ERR_NO_INTERSECTION,
 * WT_ERR_DEGENERATE, or WT_ERR_INVALID_ARG for a camera wt_projector_corners would
 * reject). All arrays are device memory; stream-ordered and capturable. Agrees with
 * wt_projector_corners up to the rounding of the device's tan, sin, cos and acos. */
#define WT_MAX_CAMERAS (1u << 24)
wt_status wt_projector_corners_batched(wt_ctx* ctx, const wt_camera* d_cameras, const float* d_amplitudes,
                                       uint32_t count, double* d_corners, int32_t* d_status, void* stream);

/* Steepness normalisation and upload value, render-engine.cpp:1071-1073:
 * Q_i = (w * A != 0) ? q / (w * A * normaliser_count) : 0, written into
 * waves[i].steepness_Q_i. normaliser_count is GerstnerWave::Count() in the
 * reference (the live wave count). */
wt_status wt_normalize_steepness(wt_gerstner_wave* waves, const float* steepness_q, uint32_t count,
                                 uint32_t normaliser_count);

/* ---------------------------------------------------------------- forward */
/* The water vertex program (water-grid.vert:114-193) over the whole
 * gridLength^2 grid: bilinear corner interpolation, Gerstner sum, heightmap
 * bump sampled at the displaced point, vertical bounce; analytic normal
 * normalize(dP/dz x dP/dx) (heightmap term included by the chain rule) flipped
 * toward the camera as at water-grid.vert:190, tangent normalize(dP/dx),
 * bitangent normalize(dP/dz). Replaces the glDrawElements of
 * render-engine.cpp:1115 plus the vertex stage it runs. d_waves: device array
 * of wave_count waves. heightmap may be NULL (no bump). */
wt_status wt_water_displace(wt_ctx* ctx, const wt_water_params* params, const wt_gerstner_wave* d_waves,
                            uint32_t wave_count, const wt_heightmap* heightmap, const wt_vertex_buffers* out,
                            void* stream);

/* The water program's remaining outputs (water-grid.vert:66-70, 141-158, 193): the
 * varyings its fragment stage reads, formed from the positions and normals
 * wt_water_displace wrote. Each array is optional (NULL skips it). */
#define WT_MAX_VARYING_VERTICES 8589934080ull /* 2 * (2^32 - 256) */
typedef struct wt_varying_buffers {
  float* clip_position;        /* float4: viewProjection * (P, 1), the vertex's gl_Position */
  float* ndc_xy_height0;       /* float2: xy / w of viewProjection * (P.x, 0, P.z, 1) */
  float* view_vector;          /* float3: cameraPosition - P (viewVecRaw, not normalised) */
  float* normal_view_only_yaw; /* float3: normalize(viewMatOnlyYaw * (N, 0)) */
  float* heightmap_sample;     /* float: the heightmap value sampled at P (heightmap_colour.r) */
} wt_varying_buffers;

/* One HBM-bound pass over vertex_count (at most WT_MAX_VARYING_VERTICES) vertices
 * (positions / normals as float3 arrays). The limit is what one launch covers: two
 * vertices per work-item and at most 2^32 - 256 work-items per dispatch.
 * view_projection = P V of the camera and view_only_yaw = V of the same camera with
 * pitch 0 (src/render-engine.cpp:278-284), both 4x4 column-major floats as the
 * reference uploads them (wt_camera_matrices gives V and P in double). params supplies
 * cameraPosition, the heightmap scales, the scroll velocity and the time; heightmap may
 * be NULL only when heightmap_sample is NULL. */
wt_status wt_vertex_varyings(wt_ctx* ctx, const wt_water_params* params, const wt_heightmap* heightmap,
                             const float view_projection[16], const float view_only_yaw[16], const float* positions,
                             const float* normals, uint64_t vertex_count, const wt_varying_buffers* out, void* stream);

/* Overrides (or with all-zero fields, restores) the launch shape of the plain
 * forward kernel. Returns WT_ERR_INVALID_ARG for unsupported values. */
wt_status wt_set_launch_config(wt_ctx* ctx, const wt_launch_config* cfg);
/* The launch shape wt_water_displace would use for this problem size. */
wt_status wt_get_launch_config(wt_ctx* ctx, uint32_t grid_length, uint32_t wave_count, wt_launch_config* cfg);

/* Times every supported launch shape (steps per lane, lanes per row and wave split,
 * i.e. vertices per workgroup, and the wave chunk) for (grid_length, wave_count) on the
 * context's device with synthetic inputs and stores the fastest in the
 * context's table (per size bucket). best_ms (may be NULL) receives its
 * per-launch time. Allocates and frees its own scratch buffers. */
wt_status wt_autotune_forward(wt_ctx* ctx, uint32_t grid_length, uint32_t wave_count, int with_heightmap,
                              wt_launch_config* best, float* best_ms);
/* wt_autotune_forward timing each shape with frames_in_flight (1..4) launches in flight
 * (frame r on stream r mod frames_in_flight, writing its own output buffers), as a
 * renderer with that many frames in flight runs it; best_ms is the time per frame.
 * Stores the winner in the same table. No reference counterpart. */
wt_status wt_autotune_forward_overlapped(wt_ctx* ctx, uint32_t grid_length, uint32_t wave_count, int with_heightmap,
                                         uint32_t frames_in_flight, wt_launch_config* best, float* best_ms);
/* The batched kernels' segment shape for patches of side ~grid_length with
 * ~wave_count waves: times each of the WT_BATCH_SHAPES shapes on a synthetic batch of
 * such patches (about 4 M vertices) and records the fastest for that (N, W) size
 * bucket (powers of two); wt_batch_create then uses it for patches in the bucket,
 * and a padded-vertex model for buckets never tuned. best_shape receives the index
 * (the shape itself through wt_batch_patch_launch_config); no reference counterpart. */
#define WT_BATCH_SHAPES 3
wt_status wt_autotune_batch(wt_ctx* ctx, uint32_t grid_length, uint32_t wave_count, int with_heightmap,
                            uint32_t* best_shape, float* best_ms);
/* Save / load the autotune tables (plain and batched) as text, one line per size
 * bucket. */
wt_status wt_autotune_save(wt_ctx* ctx, const char* path);
wt_status wt_autotune_load(wt_ctx* ctx, const char* path);

/* The water grid's index buffer (src/program.cpp:686-717) written on the device: for
 * each cell (r, c) of the gridLength x gridLength grid, row-major, the triangles
 * (r,c) (r+1,c+1) (r+1,c) and (r,c) (r,c+1) (r+1,c+1), counter-clockwise seen from
 * +y, with vertex id = row * gridLength + col + base_vertex (base_vertex = a patch's
 * vertex_offset for a batch's shared vertex arrays). d_indices holds
 * 6 (gridLength - 1)^2 uint32 in device memory. Stream-ordered and capturable. */
wt_status wt_grid_triangle_indices(wt_ctx* ctx, uint32_t grid_length, uint32_t base_vertex, uint32_t* d_indices,
                                   void* stream);

/* ---------------------------------------------------------------- fragment stage */
/* One fragment of the water grid as the rasteriser hands it to the fragment shader
 * (assets/shaders/water-grid.frag:57-61 plus gl_FragCoord.z): the varyings
 * wt_vertex_varyings writes, interpolated to the pixel. 48 bytes; a fragment buffer
 * holds one per pixel, rows bottom to top as GL numbers window rows, 16-byte aligned. */
typedef struct wt_water_fragment {
  float normal[3];               /* `normal`, interpolated (not renormalised, as the shader reads it) */
  float view_vector[3];          /* viewVecRaw */
  float normal_view_only_yaw[3]; /* normalVecInViewSpaceOnlyYaw */
  float ndc_xy_height0[2];       /* xyPositionNDCSpaceHeight0 */
  float frag_depth;              /* gl_FragCoord.z, window-space depth in [0, 1] */
} wt_water_fragment;

/* The fragment shader's uniforms (water-grid.frag:34-55), GLSL names, as
 * RenderEngine::render sets them (src/render-engine.cpp:1060-1098; defaults in
 * src/render-engine.h:213-235). */
typedef struct wt_shade_params {
  float fogColourFarAtCurrentTime[4];
  float fogDepthRadiusFar;
  float fogDepthRadiusNear;
  float softEdgesDeltaDepthThreshold;
  float sunPosition[3];
  float sunShininess; /* >= 0 */
  float sunStrength;
  float tintDeltaDepthThreshold;
  float viewportWidthHeight[2];
  float waterClarity;
  float zFar; /* zFar > zNear > 0 */
  float zNear;
} wt_shade_params;

/* An RGBA8 image in device memory (GL_RGBA / GL_UNSIGNED_BYTE, 4-byte aligned), rows
 * bottom to top; sides 1..32768. */
typedef struct wt_rgba8_image {
  const uint8_t* texels;
  uint32_t width, height;
} wt_rgba8_image;

/* The shader's samplers with the reference's filtering (src/render-engine.cpp:90-188):
 * depthTexture2D (one float per texel, GL_NEAREST), localReflectionsTexture2D and
 * localRefractionsTexture2D (GL_LINEAR), skybox (six skybox_length^2 RGBA8 faces back
 * to back in GL target order +X, -X, +Y, -Y, +Z, -Z, GL_LINEAR per face), all
 * CLAMP_TO_EDGE. */
typedef struct wt_shade_textures {
  const float* depth;
  uint32_t depth_width, depth_height;
  wt_rgba8_image local_reflections;
  wt_rgba8_image local_refractions;
  const uint8_t* skybox;
  uint32_t skybox_length;
} wt_shade_textures;

/* The water program's fragment stage (water-grid.frag:70-162) as a deferred pass over
 * a width x height fragment buffer (sides 1..32768): for every pixel whose coverage
 * byte is nonzero (coverage NULL: every pixel), the fragment colour RGBA (float4, 16-byte
 * aligned colour buffer) with gl_FragCoord.xy at the pixel centre. blend = 0 writes the
 * colour; blend = 1 blends it over the colour buffer's value as the reference's draw
 * does (glBlendFunc(GL_SRC_ALPHA, GL_ONE_MINUS_SRC_ALPHA), src/render-engine.cpp:
 * 302-303, on all four channels of a floating-point buffer). Uncovered pixels are not
 * touched. Replaces the fragment half of the glDrawElements at
 * src/render-engine.cpp:1115; the rasterisation that produces the fragments stays
 * with the renderer. Stream-ordered and capturable. */
wt_status wt_water_shade(wt_ctx* ctx, const wt_shade_params* params, const wt_shade_textures* textures,
                         const wt_water_fragment* fragments, const uint8_t* coverage, uint32_t width,
                         uint32_t height, float* colour, int blend, void* stream);

/* ---------------------------------------------------------------- batched forward */
/* One ocean patch of a batch: its own projected grid (corners from
 * wt_projector_corners of its own camera), time, scales and a run of the batch's
 * wave array. Outputs go to [vertex_offset, vertex_offset + gridLength^2) of the
 * batch's output arrays. The reference draws one grid per frame
 * (src/render-engine.cpp:1115); a batch replaces many such draws. */
typedef struct wt_patch {
  double corners[4][4];
  float cameraPosition[3];
  uint32_t gridLength;
  uint32_t wave_offset;
  uint32_t wave_count;
  uint64_t vertex_offset;
  double waveAnimationTimeInSeconds;
  float heightmapSampleScale;
  float heightmapDisplacementScale;
  float verticalBounceWaveDisplacement;
  float heightmapScrollVelocity[2];
  float reserved;
} wt_patch;

/* Validates the patches, uploads them, and builds the work list (units of one
 * workgroup's vertices of one patch, grouped by segment shape, most expensive first). */
wt_status wt_batch_create(wt_ctx* ctx, const wt_patch* patches, uint32_t count, wt_batch** out);
void wt_batch_destroy(wt_batch* batch);
/* Total vertices of the batch (the size the output arrays need, in float3). */
wt_status wt_batch_vertex_count(const wt_batch* batch, uint64_t* vertices);
/* The segment shape the batch runs patch `patch` with (chosen per patch at creation
 * to minimise padded vertices; no reference counterpart). wt_set_launch_config with the
 * same shape makes wt_water_displace on that patch alone produce identical bits. */
wt_status wt_batch_patch_launch_config(const wt_batch* batch, uint32_t patch, wt_launch_config* cfg);
/* Re-aims every patch of a batch from its own camera on the device: patch i takes
 * the corners of d_cameras[i] with slab half-height d_amplitudes[i] (as
 * wt_projector_corners_batched) and cameraPosition = its position. A patch whose
 * camera has no corners (status other than WT_OK, written to d_status[i] when
 * d_status is not NULL) is skipped by later wt_water_displace_batched calls, which
 * leave its outputs untouched, as the reference skips the water draw
 * (src/render-engine.cpp:894), until a later call gives it corners again. Arrays
 * hold one entry per patch in device memory; stream-ordered and capturable, so a
 * captured frame loop moves cameras without a host round trip. */
wt_status wt_batch_set_cameras(wt_ctx* ctx, wt_batch* batch, const wt_camera* d_cameras, const float* d_amplitudes,
                               int32_t* d_status, void* stream);
/* Every patch's vertex program at time + time_offset, in one launch per segment
 * shape (one workgroup per unit, dispatched largest first). d_waves holds total_waves waves that the patches index into.
 * Replayable from a captured graph (no device state changes between calls). */
wt_status wt_water_displace_batched(wt_ctx* ctx, wt_batch* batch, const wt_gerstner_wave* d_waves, uint32_t total_waves,
                                    const wt_heightmap* heightmap, const wt_vertex_buffers* out, double time_offset,
                                    void* stream);

/* ---------------------------------------------------------------- graphs */
/* Stream capture of any sequence of wt_* calls into a HIP graph, so a frame
 * loop replays with one launch (the reference's per-frame render loop,
 * src/program.cpp:83-99). */
wt_status wt_graph_begin(wt_ctx* ctx, void* stream);
wt_status wt_graph_end(wt_ctx* ctx, void* stream, wt_graph** out);
wt_status wt_graph_launch(wt_graph* graph, void* stream);
void wt_graph_destroy(wt_graph* graph);

/* ---------------------------------------------------------------- heightfield fit */
/* Fitting a wave set to observed displacements (no reference counterpart: the
 * reference only evaluates waves, water-grid.vert:74-85). Model per wave
 * {A, L, alpha, Q, phi0}: w = 2 pi / L, omega = sqrt(g w), D = (cos alpha,
 * sin alpha), theta = w D.x + omega t + phi0 (the reference's phase,
 * water-grid.vert:78), displacement (Q A Dx cos, A sin, Q A Dz cos)
 * (water-grid.vert:79-81). Loss = 1/(2 S V) sum |displacement - target|^2 over
 * an n x n world grid and S snapshots. */
typedef struct wt_fit_problem {
  uint32_t grid_side;  /* n grid points per side, 1..8192 */
  uint32_t snapshots;  /* S, 1..32, at t0 + s dt */
  uint32_t wave_count; /* W, 1..128 */
  uint32_t reserved;
  double spacing;      /* metres between grid points */
  double x0, z0;       /* world position of grid point (0, 0); point (i, j) = (x0 + i spacing, z0 + j spacing) */
  double t0, dt;
  double gravity;
  const float* target; /* device [S][n*n][3] displacement (dx, y, dz), row-major points */
  float* params;       /* device [W][5] {A, L, alpha, Q, phi0}; Adam updates it in place */
} wt_fit_problem;

typedef struct wt_adam_config {
  float lr;
  float beta1;
  float beta2;
  float eps;
} wt_adam_config;

/* A batch of fitting problems of any sizes; every call below covers the whole batch
 * in one pair of kernel launches (adjoint, then reduce + Adam). Allocates the Adam
 * moments and per-block partial sums once. */
wt_status wt_fit_create(wt_ctx* ctx, const wt_fit_problem* problems, uint32_t count, wt_fit** out);
void wt_fit_destroy(wt_fit* fit);
/* Zeroes the Adam moments and step counters (a kernel, safe inside a captured graph). */
wt_status wt_fit_reset(wt_fit* fit, void* stream);
/* `iterations` Adam steps on every problem. The step counter lives on the device, so
 * a captured sequence of iterations replays correctly. d_losses (may be NULL): device
 * array of `count` floats, receives each problem's loss before its last step. */
wt_status wt_fit_iterate(wt_fit* fit, const wt_adam_config* adam, uint32_t iterations, float* d_losses, void* stream);
/* Loss and gradient at the current params: d_grad receives the problems' W*5
 * gradients concatenated in batch order, d_losses (may be NULL) one loss each. */
wt_status wt_fit_loss_grad(wt_fit* fit, float* d_grad, float* d_losses, void* stream);
/* wt_fit_loss_grad for one problem without a wt_fit and without allocating: no
 * counterpart in the reference; for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * descriptor, maps and per-block partial sums live in a fixed device array of
 * WT_FIT_STATIC_SCRATCH_BYTES per GPU, held until the process ends; a problem needing
 * more (about (5 W + 1) x ceil(n^2 / 256) x 4 bytes of partials, half the blocks for
 * S <= 8) is refused with WT_ERR_INVALID_ARG. d_grad receives W*5 floats, d_loss (may be
 * NULL) one. Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (on different streams or host threads) race and must be ordered by the caller. */
#define WT_FIT_STATIC_SCRATCH_BYTES (16u << 20)
wt_status wt_fit_loss_grad_static(wt_ctx* ctx, const wt_fit_problem* problem, float* d_grad, float* d_loss,
                                  void* stream);
/* Writes the displacement of problem->params (taken as the hidden wave set) at every
 * grid point and snapshot into d_target, plus N(0, sigma) noise per component from a
 * counter-based generator keyed by seed (problem->target is ignored). */
wt_status wt_fit_synthesize(wt_ctx* ctx, const wt_fit_problem* problem, float sigma, uint64_t seed, float* d_target,
                            void* stream);
/* Phase-storage policy of the adjoint: the bytes a stored (cos, sin) tensor of
 * S x n^2 x W would take, and whether the kernels recompute the phases instead
 * (always 1 in this build: recomputing by rotation is cheaper than the HBM round
 * trip at every size). */
wt_status wt_fit_phase_plan(const wt_fit_problem* problem, uint64_t budget_bytes, uint64_t* phase_tensor_bytes,
                            int* recompute);

#ifdef __cplusplus
}
#endif

#endif /* WT_H_ */
