/* stipple_amd.h -- the C-ABI of the MI355X Riemersma hot path.
 *
 * Plain C, plain pointers and sizes; no C++ or torch types cross this line.
 * Every function returns 0 (SA_OK) on success or a nonzero SA_E* code, and
 * writes a human-readable message into `err` (when non-null) on failure only
 * -- the same success-is-empty convention as the reference's engines
 * (src/rd_cuda_stub.cpp:60-74: an engine never reports success silently).
 *
 * Device pointers are HIP device pointers on the current device.  Calls that
 * take a `stream` (a hipStream_t, or NULL for the null stream) are
 * stream-ordered: they enqueue work and return; results are ready when the
 * stream reaches that point.  Calls without a stream are synchronous.
 *
 * The C++ drop-in layer (include/rd_riemersma.h, include/rd_octree.h,
 * include/rd_plugin.h in this repository) is built on these entry points; each
 * entry point names the reference interface it replaces.
 */
#ifndef STIPPLE_AMD_H_
#define STIPPLE_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SA_OK = 0,
  SA_EINVAL = 1,       /* bad argument */
  SA_EREFUSED = 2,     /* an input the reference engines also refuse */
  SA_EDEVICE = 3,      /* HIP runtime error, or no device */
  SA_ENOMEM = 4,       /* device allocation failed */
  SA_EUNSUPPORTED = 5  /* valid input this build does not handle */
};

/* Pixel formats: interleaved, row-major, dense (width*height pixels), except
 * the planar YUV wire formats of the video path, which only the block engine
 * (sa_walk_blocks, sa_chain_create_blocks) and the palette build
 * (sa_palette_build, _batch) read, and only the block engine writes:
 *   SA_PIX_YUV444P8: Y, Cb, Cr planes of width*height bytes each.  Input
 *     widens through swscale's integer matrix (SwsYuvToRgb16,
 *     include/rd_types.h:75-97; BlockOptions::InMode::PlanarYuv444 and
 *     PlanarYuv444Prepass, include/rd_riemersma.h:325-443) with alpha 65535;
 *     output is the limited-range BT.601 matrix of the chosen entry
 *     (BlockOptions::emit_yuv444; src/rd_riemersma_cpu.cpp:718-730).
 *   SA_PIX_YUV420P8: Y (width*height), then Cb and Cr at (width/2)*(height/2);
 *     input only, even width and height; chroma is upsampled bilinearly
 *     (InMode::PlanarYuv420, src/rd_blocks_cuda.cu:651-740). */
enum {
  SA_PIX_RGBAF32 = 0, /* RgbaF: 4 x float Q16, the reference's in-memory layout */
  SA_PIX_RGB8 = 1,    /* 8-bit samples widen x257 on input */
  SA_PIX_RGBA8 = 2,
  SA_PIX_RGB16 = 3,   /* 16-bit samples are Q16 as-is */
  SA_PIX_RGBA16 = 4,  /* rgba64le */
  SA_PIX_INDEX8 = 5,  /* output only: palette index per pixel */
  SA_PIX_INDEX16 = 6,
  SA_PIX_YUV444P8 = 7,
  SA_PIX_YUV420P8 = 8
};

typedef struct sa_palette_s* sa_palette;

/* ---- runtime ----------------------------------------------------------- */

/* Replaces rd::CudaAvailable (include/rd_riemersma.h:284): 1 when a gfx950
 * device is usable. */
int sa_available(void);

/* Selects the device and warms the runtime.  Replaces the implicit
 * cudaGetDevice of src/rd_riemersma_cuda.cu:561-562. */
int sa_init(int device, char* err, size_t errlen);

/* Device name (the `device_name` out-parameter of the reference engines,
 * include/rd_riemersma.h:214-217). */
int sa_device_name(char* out, size_t len);

/* ---- palette + octree --------------------------------------------------- */

/* Uploads a palette and the octree that its lookups search.  Replaces the
 * host-side UploadTree / UploadPalette / BuildFlatSearch of
 * include/rd_cuda_common.cuh:290-360,479-522.  `entries` holds count*4
 * doubles (r,g,b,a: rd::Palette::entries); the tree arrays are rd::QNode
 * fields, `nodes` long (child: 16 per node). */
int sa_palette_create(const double* entries, int count, int assoc, const int32_t* child16,
                      const int32_t* parent, const uint32_t* color_number,
                      const uint64_t* number_unique, int64_t nodes, sa_palette* out,
                      char* err, size_t errlen);
/* Releases a palette.  No work queued after the call may use it; work queued
 * before it may still be running.  The call does not wait for the device: the
 * palette's memory goes back to the runtime with those of later destroys,
 * after one device synchronisation once they hold more than 256 MiB. */
int sa_palette_destroy(sa_palette p);
int sa_palette_count(sa_palette p);

/* ColorTree::Build on the device (classify, reduce, colormap).  Replaces
 * rd::ColorTree::Build (include/rd_octree.h:62-63, src/rd_octree.cpp:219-409);
 * bit-exact with it: same colormap, same node numbering, the same doubles in
 * every node but one.  The root's quantize_error (a sum over every
 * classification step) is summed in parallel within a rounding bound; its
 * value can affect the reduction only if it comes within 1 of another node's
 * error, and the serial chain runs exactly when the bound cannot rule that
 * out.  So every decision and the colormap are exact, and the root's stored
 * double may differ in its last bits (sa_palette_info's exact_root says
 * whether it was summed serially).  `src` is a device image in src_fmt
 * (planar YUV included: the
 * tree of the frame as the reader widens it).  Synchronises `stream`
 * (the build reads back a few scalars).  With keep_tree the node arrays stay
 * on the device for sa_palette_tree.  A tree whose node total could pass
 * MaxQNodes (266817) during classification -- palettes of about a thousand
 * colours or more on busy images -- is classified again row by row, as
 * quantize.c does: PruneLevel at the start of the row after the live count
 * passed the limit, later rows one level shallower (src/rd_octree.cpp:366-369,
 * :380-383).  Its every node, the pruned ones and the root included, is then
 * the reference's exactly; such a build takes a few hundred milliseconds per
 * megapixel. */
int sa_palette_build(const void* src, int src_fmt, int64_t width, int64_t height, int colors,
                     int assoc, int gray, int keep_tree, void* stream, sa_palette* out,
                     char* err, size_t errlen);

/* One image of a batched build: the arguments of sa_palette_build. */
typedef struct {
  const void* src;
  int src_fmt;
  int64_t width, height;
  int colors;
  int assoc, gray, keep_tree;
} sa_build;

/* ColorTree::Build for n images at once (rd::ColorTree::Build,
 * include/rd_octree.h:62-63, once per image): every kernel is launched once
 * for the whole batch and `stream` is synchronised four times per batch, not
 * per image.  out[i] is exactly the palette sa_palette_build would return
 * for items[i], or NULL with status[i] (status may be NULL) set to its
 * error.  Returns SA_OK when every item built, else the first failing
 * item's code, its message in err. */
int sa_palette_build_batch(const sa_build* items, int n, void* stream, sa_palette* out,
                           int* status, char* err, size_t errlen);

/* ---- video palette: montage of sampled frames -----------------------------
 * VideoBuildPalette's montage path (src/rd_video.cpp:1816, :2080-2700): the
 * `n` sampled frames (device pointers, one geometry width x height, src_fmt
 * SA_PIX_RGB8 / RGBA8 / RGB16 / RGBA16 -- the decoder's rgb24 / rgba /
 * rgba64le) are placed in sample order in the cells of a grid `cols` cells
 * wide (0: ceil(sqrt(want)), the reference's default; `want` cells in all,
 * rows = ceil(want / cols)).  A cell is the whole frame (tile <= 0) or a
 * tile x tile nearest-neighbour sample of it; cells past the last sample repeat
 * it.  8-bit samples widen x257, 3-channel sources get alpha 65535.
 *
 * sa_video_montage writes that montage (device, *mw x *mh RGBA16, exactly the
 * reference's RgbaF montage) and sets *gray when every pixel has r == g == b;
 * with montage == NULL it only reports the size.  Synchronises `stream` when
 * gray is non-null.  Montages above 2^28 pixels are refused (the reference's
 * own montage cap is 64 Mpixel: reduce `want` to fit before sampling). */
int sa_video_montage(const void* const* frames, int n, int src_fmt, int64_t width,
                     int64_t height, int want, int tile, int cols, void* montage, int64_t* mw,
                     int64_t* mh, int* gray, void* stream, char* err, size_t errlen);

/* The palette of that montage: ColorTree::Build at `colors`, alpha not
 * associated, greyscale as detected on the montage -- the tree that
 * ImBuildPaletteFromPixels builds (src/rd_im.cpp:535-656); bit-exact with
 * rd::ColorTree::Build on the montage.  (HarvestColormap's removal of exact
 * duplicate colormap entries, src/rd_im.cpp:555-571, is applied by the C++
 * layer's ImBuildPaletteFromPixels, not here.) */
int sa_video_palette(const void* const* frames, int n, int src_fmt, int64_t width,
                     int64_t height, int want, int tile, int cols, int colors, int keep_tree,
                     void* stream, sa_palette* out, char* err, size_t errlen);

/* ---- palette text files -----------------------------------------------------
 * The .txt / .text form of ImWritePalette / ImReadPalette (include/rd_im.h:
 * 128-186; src/rd_im.cpp:994-1024, :1175-1330).  Host only.  Write: `count`
 * entries (4 doubles each) as "index r g b a RRGGBB" rows at %.6f after two
 * comment lines.  Read: `entries` needs room for 65536 x 4 doubles; rows may
 * omit alpha (opaque); indices must be contiguous from 0; a positive
 * expect_colors must equal the file's count (SA_EREFUSED otherwise); r, g, b
 * come back narrowed through float and the entries sorted by (r, g, b, a);
 * *assoc is 1 when some alpha is not 65535.  The reference's importer then
 * builds its tree from those entries as a count x 1 strip at `count` colours
 * and adopts the tree's colormap (the C++ layer's ImReadPalette does that on
 * the device).  Other file types (image palettes) return SA_EUNSUPPORTED. */
int sa_palette_write_txt(const double* entries, int count, const char* path, char* err,
                         size_t errlen);
int sa_palette_read_txt(const char* path, int expect_colors, double* entries, int* count,
                        int* assoc, char* err, size_t errlen);

/* Palette facts: colours, association, tree depth, node count (ColorTree::
 * color_count / depth / node_count, include/rd_octree.h:65-99) and whether the
 * root's quantize_error was summed serially (1) or bounded (0). */
int sa_palette_info(sa_palette p, int* count, int* assoc, int* depth, int64_t* nodes,
                    int* exact_root);

/* The colormap, count*4 doubles (ColorTree::colormap, include/rd_octree.h:66). */
int sa_palette_entries(sa_palette p, double* out, char* err, size_t errlen);

/* The node arrays of a tree built with keep_tree (ColorTree::nodes,
 * include/rd_octree.h:67); each array `nodes` long, child 16 per node. */
int sa_palette_tree(sa_palette p, int32_t* child16, int32_t* parent, uint32_t* color_number,
                    uint64_t* number_unique, double* total4, double* qe, uint32_t* id,
                    uint32_t* level, char* err, size_t errlen);

/* ---- curve --------------------------------------------------------------- */

/* Hilbert visit order of a w x h image, generated on the device: the
 * w*h in-bounds pixel ids in the order the sequential walk visits them.
 * Replaces rd::BuildCurveIndex / BlockCurveOrder (include/rd_riemersma.h:
 * 178-179, 278-279); the block curve is this order without its last entry
 * when w is a power of two >= h, plus the origin.  `order` is a device array
 * of w*h int32. */
int sa_curve_order(int64_t w, int64_t h, int32_t* order, void* stream, char* err,
                   size_t errlen);

/* ---- engines -------------------------------------------------------------- */

typedef struct {
  const void* src; /* device pixels, src_fmt */
  void* dst;       /* device pixels, dst_fmt; may equal src for the sequential engine */
  int64_t width;
  int64_t height;
  int32_t src_fmt;
  int32_t dst_fmt;
  sa_palette palette;
} sa_image;

typedef struct {
  double diffusion; /* dither:diffusion-amount (rd::DitherParams::diffusion), 1.0 */
  int32_t use_cache;/* quantize.c's memo; 1 reproduces ImageMagick */
  int32_t memo_bits;/* LDS memo line slots per wavefront = 2^memo_bits (4..10); 0 = as many
                       as the LDS left beside the palette holds (byte entries for palettes
                       of at most 255 colours, so twice the lines) */
  int32_t waves;    /* persistent wavefronts, 0 = auto */
  int32_t reserved;
} sa_seq_options;

/* The exact sequential walk over a batch of independent images, one
 * wavefront per image, pulled from a device queue largest-first.  Replaces
 * rd::RiemersmaWalkCuda (include/rd_riemersma.h:214-217, whose `frames`
 * copies become arbitrary (image, palette) items).  Bit-exact per image with
 * rd::RiemersmaWalkCpu. */
int sa_walk_seq_batch(const sa_image* images, int n, const sa_seq_options* opt, void* stream,
                      char* err, size_t errlen);

/* ---- resumable sequential walks (continuous batching) ------------------
 * New, with no counterpart in the reference (its RiemersmaWalkCuda walks one
 * batch per call).  A job holds images that join while others are on the GPU:
 * each step advances every unfinished image by a slice of visits in one launch
 * and packs the ones that finish.  A sliced walk saves its queue and memo
 * cache between launches, so every image is bit-exact with one uninterrupted
 * RiemersmaWalkCpu walk, whatever it shared the device with.  Progress is
 * deterministic (min(n, t + slice) per step), so the host knows which images
 * finish in a step without reading the device. */
typedef struct sa_walk_job_s* sa_walk_job;

int sa_walk_job_create(const sa_seq_options* opt, sa_walk_job* out, char* err, size_t errlen);

/* Adds `n` images (palettes must outlive their walks).  Stream-ordered: the
 * inputs are gathered and the memo tables cleared on `stream`; use the step
 * stream, or order the two yourself.  ids[i] receives each image's id. */
int sa_walk_job_add(sa_walk_job job, const sa_image* images, int n, int64_t* ids, void* stream,
                    char* err, size_t errlen);

/* One step on `stream`: every unfinished image advances by `visits` visits
 * (rounded up to a multiple of 48; <= 0 runs them to the end), then the images
 * that finished are packed into their dst.  done[0..min(cap, *n_done)) gets
 * their ids; *active the images still unfinished.  Outputs of done images are
 * ready when `stream` reaches this point. */
int sa_walk_job_step(sa_walk_job job, int64_t visits, void* stream, int64_t* done, int cap,
                     int* n_done, int* active, char* err, size_t errlen);

/* The same step with a time budget instead of a visit count: every image
 * walks until `budget_us` microseconds after its wavefront picked it up
 * (checked every 48 visits), so images progress at their own pace and none
 * waits for the slowest.  Progress is read back: this synchronises `stream`. */
int sa_walk_job_step_timed(sa_walk_job job, double budget_us, void* stream, int64_t* done,
                           int cap, int* n_done, int* active, char* err, size_t errlen);

/* The timed step with the host out of the device's way: launches this step,
 * then waits for the previous step's progress (copied back asynchronously) and
 * packs the images that finished in it, ordered after this step.  done[] and
 * *active describe what is known after the previous step, so an image is
 * reported one call after it finishes; call until *active is 0.  Outputs of
 * done images are ready when `stream` reaches this point. */
int sa_walk_job_step_overlapped(sa_walk_job job, double budget_us, void* stream, int64_t* done,
                                int cap, int* n_done, int* active, char* err, size_t errlen);

/* Frees the job (waits for the device if images are unfinished). */
int sa_walk_job_destroy(sa_walk_job job);

/* Counters of the sequential walks since the last call, then reset:
 * out[0] memo misses (full octree searches), out[1] LDS memo line fills,
 * out[2] line fills that read HBM, out[3] reserved.  The analogue of the
 * reference's RD_CC_STATS (include/rd_cuda_common.cuh:210-229). */
int sa_seq_stats(uint64_t* out, char* err, size_t errlen);

/* The block-parallel walk over `n` frames of one geometry (each frame may
 * carry its own palette).  Replaces rd::RiemersmaBlocksCuda
 * (include/rd_riemersma.h:450-464); bit-exact with rd::RiemersmaBlocksCpu at
 * the same `block`.  dst must not alias src.  Refuses block < 16 and palettes
 * above 256 entries, as the reference does. */
int sa_walk_blocks(const sa_image* frames, int n, int block, double diffusion, void* stream,
                   char* err, size_t errlen);

/* sa_walk_blocks without allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments,
 * results and refusals as sa_walk_blocks; the walk's scratch (frame table,
 * weights, curve) lives in one fixed 40 MiB device array per device, bound by
 * sa_init, so the call makes no device or pinned-host allocation.  Limits:
 * sides up to 4096, width*height up to 10469376 (a 3840x2160 frame fits), at
 * most 1024 frames; beyond them it returns SA_EUNSUPPORTED.  Requires sa_init
 * on the current device (else SA_EINVAL).  Every call on a device shares that
 * array: calls that overlap on one device race. */
int sa_walk_blocks_noalloc(const sa_image* frames, int n, int block, double diffusion,
                           void* stream, char* err, size_t errlen);

/* The threshold-screen dithers of the plugin seam, one thread per pixel, over
 * frames of one geometry (each with its own palette, any entry count):
 *   SA_SCREEN_BAYER8     "bayer-ordered" (src/rd_dithers_ordered.cc:147-168)
 *   SA_SCREEN_VOID64     "void-and-cluster" (:220-330, :447-482)
 *   SA_SCREEN_VOID32     "void-and-cluster-fast" (:484-486)
 *   SA_SCREEN_CLUSTERED8 "clustered-dot" (src/rd_dithers.cc:294-358)
 * Each replaces that DitherAlgorithm's run (include/rd_plugin.h:132-136) on
 * device frames: the chosen entry's r, g, b with the source's alpha
 * (WriteChosen), bit-exact with the plugin.  `diffusion` is DitherJob's.
 * dst may equal src when the formats match.  Refusals: a palette without
 * entries ("<name>: no palette") and an empty frame ("<name>: empty frame"),
 * as Prepare (src/rd_dither_common.h:108-120). */
enum { SA_SCREEN_BAYER8 = 0, SA_SCREEN_VOID64 = 1, SA_SCREEN_VOID32 = 2, SA_SCREEN_CLUSTERED8 = 3 };
int sa_dither_screen(const sa_image* frames, int n, int screen, double diffusion, void* stream,
                     char* err, size_t errlen);

/* The error-diffusion dithers of the plugin seam, over frames of one geometry
 * (each with its own palette, any entry count), one frame per wavefront:
 *   SA_DIFFUSE_FLOYD    "floyd-steinberg" (src/rd_dithers.cc:74-110)
 *   SA_DIFFUSE_ATKINSON "atkinson"        (:112-148)
 *   SA_DIFFUSE_JARVIS   "jarvis"          (:150-220)
 * Each replaces that DitherAlgorithm's run (include/rd_plugin.h:132-136) on
 * device frames, bit-exact with the plugin: the same float arithmetic in the
 * same order, the chosen entry's r, g, b with the source's alpha.  `diffusion`
 * is DitherJob's.  dst may equal src when the formats match.  Frames keep
 * their error rows in LDS: widths above about 6800 (floyd-steinberg,
 * atkinson) or 4500 (jarvis) return SA_EUNSUPPORTED.  Refusals as
 * sa_dither_screen ("<name>: no palette", "<name>: empty frame"). */
enum { SA_DIFFUSE_FLOYD = 0, SA_DIFFUSE_ATKINSON = 1, SA_DIFFUSE_JARVIS = 2 };
int sa_dither_diffuse(const sa_image* frames, int n, int kind, double diffusion, void* stream,
                      char* err, size_t errlen);

/* The Jacobi approximation: `iterations` fixed-point sweeps of
 * t = pal[q] + F (x - pal[q]) along the block curve, F = (I-H)^-1 truncated
 * to `taps`+1 terms, then an owner-filtered scatter.  Replaces
 * rd::RiemersmaApproxCuda (include/rd_riemersma.h:296-307; ApproxOptions
 * iterations / taps / fp64 / use_tree).  Frames share one geometry and may
 * carry their own palettes; dst must not alias src.  Refuses palettes above
 * 256 entries and taps outside 1..4096, as the reference does. */
int sa_walk_approx(const sa_image* frames, int n, int iterations, int taps, int fp64,
                   int use_tree, double diffusion, void* stream, char* err, size_t errlen);

/* The inverse filter g[0..taps] of the Jacobi engine, (I - H)^-1 truncated
 * to taps+1 terms, computed on the host with the reference's weight
 * expression.  Replaces rd::ApproxInverseFilter (include/rd_riemersma.h:309;
 * src/rd_approx_cuda.cu:73-88).  `g` holds taps+1 doubles; taps in 1..4096. */
int sa_inverse_filter(int taps, double diffusion, double* g, char* err, size_t errlen);

/* ---- frame chains (hipGraph) --------------------------------------------
 * sa_walk_blocks or sa_walk_approx for a fixed batch of frame buffers,
 * captured once into a hipGraph that owns its scratch, its copy of the curve
 * and its weights.  A video loop that refills the same src buffers (and keeps
 * the palettes) replays the whole batch with one sa_chain_run: the engine's
 * launches without their per-call validation, uploads and allocations.  Each
 * replay computes exactly what the eager call computes on the buffers'
 * current contents.  New, with no counterpart in the reference (whose
 * RiemersmaBlocksCuda / RiemersmaApproxCuda, include/rd_riemersma.h:450-464,
 * :296-307, are one eager call per batch). */
typedef struct sa_chain_s* sa_chain;

/* Same arguments, validation and refusals as sa_walk_blocks.  `stream` orders
 * the set-up copies; the graph itself is captured on a private stream. */
int sa_chain_create_blocks(const sa_image* frames, int n, int block, double diffusion,
                           void* stream, sa_chain* out, char* err, size_t errlen);

/* Same arguments, validation and refusals as sa_walk_approx. */
int sa_chain_create_approx(const sa_image* frames, int n, int iterations, int taps, int fp64,
                           int use_tree, double diffusion, void* stream, sa_chain* out,
                           char* err, size_t errlen);

/* Replays the chain on `stream` (stream-ordered; the frames' buffers and
 * palettes must be alive and not written by other work meanwhile). */
int sa_chain_run(sa_chain chain, void* stream, char* err, size_t errlen);

/* Frees the chain (waits for the device). */
int sa_chain_destroy(sa_chain chain);

/* ---- tuning knobs -------------------------------------------------------
 * Launch shapes and code paths an autotuner or a test may set
 * (stipple_amd/tune.py measures the first three); no knob changes a palette or
 * an output.  New, with no counterpart in the reference.
 *   "blocks.threads"  block walk: Hilbert blocks per workgroup (64..512; 256)
 *   "approx.per"      Jacobi sweep: outputs per thread (2, 4, 8; 4)
 *   "seq.memo_bits"   sequential walk: a cap of 2^value LDS memo slots (4..10;
 *                     0 = no cap: the most that fit, which leaves one image per CU)
 *   "octree.cells"    palette build: 1 places loop 2's pairs by a counting sort
 *                     over cells when the cube depth leaves at most 4680 cells,
 *                     0 sends every pair through the node hash (1)
 *   "octree.exact_root" palette build: 1 sums the root's quantize_error serially
 *                     always (the palette is the same either way; only the
 *                     tree's root value becomes exact), 0 only when it could
 *                     matter (0)
 * Calls made after a set use the new value; chains keep the value they were
 * created with. */
int sa_tune_set(const char* knob, int value, char* err, size_t errlen);
int sa_tune_get(const char* knob, int* value);

/* ---- kernel timing --------------------------------------------------------- */

/* Brackets every engine kernel launch with HIP events on its own stream while
 * enabled -- the RD_KERNEL_TIMING diagnostic of the reference
 * (src/rd_blocks_cuda.cu:820-837).  Kernel names: seq_gather, seq_walk,
 * seq_pack, blocks_walk, approx_sweep, octree. */
int sa_timing_enable(int on);

/* Sums the recorded durations of kernel `name` (all kernels when NULL),
 * waiting for them to finish, and forgets them. */
int sa_timing_collect(const char* name, double* total_ms, int* launches, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* STIPPLE_AMD_H_ */
