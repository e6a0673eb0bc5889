/* oracle.h -- CPU restatement of the reference Riemersma hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (stipple_amd/) links,
 * loads or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker or as the
 * timed CPU baseline -- never as the thing measured or shipped.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Arithmetic order and float width follow
 * the reference exactly; the library is built with -O2 -ffp-contract=off and
 * no fast-math so every double/float operation rounds once, as there.
 *
 * Parity status of each function (how it is pinned) is listed in oracle.cpp's
 * header comment.
 */
#ifndef STIPPLE_ORACLE_H_
#define STIPPLE_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- numeric contract (include/rd_types.h) ---------------------------- */
void orc_error_weights(double out[16]);                   /* rd_types.h:169-176 */
void orc_scaled_weights(double diffusion, double out[16]); /* rd_riemersma_cpu.cpp:370-373 */
unsigned orc_scale_char(float q);                         /* rd_types.h:133-138 */
float orc_clamp_pixel(double v);                          /* rd_types.h:39-43 */

/* ---- curve (include/rd_riemersma.h:88-157, src/rd_riemersma_cpu.cpp) ---- */
int orc_curve_level(int64_t w, int64_t h);                /* rd_riemersma.h:88-94 */
void orc_hilbert_point(int64_t n, int level, int* x, int* y); /* rd_riemersma_cpu.cpp:167-189 */
/* The explicit recursion of IM's Riemersma() (rd_riemersma_cpu.cpp:126-153):
 * writes the 4^level leaf positions followed by the trailing ForgetGravity
 * visit.  Returns the number of entries (4^level, or 1 at level 0). */
int64_t orc_recursive_walk(int level, int* xs, int* ys);
/* In-bounds pixel ids of the sequential walk in visit order (w*h of them). */
int64_t orc_seq_order(int64_t w, int64_t h, int32_t* order);
/* Block-engine curve: in-bounds of n in [0,4^L-1) plus a trailing origin, and
 * the last-writer owner map (rd_riemersma_cpu.cpp:269-356).  order must hold
 * w*h+1 entries, owner w*h.  Returns the curve length. */
int64_t orc_block_order(int64_t w, int64_t h, int32_t* order, int32_t* owner);

/* ---- colour octree (src/rd_octree.cpp) -------------------------------- */
typedef struct OrcTree OrcTree;
/* rgba: w*h*4 floats, row-major, Q16.  rd_octree.cpp:219-409 */
OrcTree* orc_tree_build(const float* rgba, int64_t w, int64_t h, int max_colors,
                        int assoc, int gray);
void orc_tree_free(OrcTree* t);
int orc_tree_color_count(const OrcTree* t);
int64_t orc_tree_node_count(const OrcTree* t);
int orc_tree_depth(const OrcTree* t);
int orc_tree_assoc(const OrcTree* t);
/* colormap as count*4 doubles (r,g,b,a). */
void orc_tree_colormap(const OrcTree* t, double* out);
/* node arrays, each node_count long (child: 16 per node). */
void orc_tree_nodes(const OrcTree* t, int32_t* child16, int32_t* parent,
                    uint32_t* color_number, uint64_t* number_unique,
                    double* total4, double* qe, uint32_t* id, uint32_t* level);
/* Descent + ClosestColor for one clamped target (rd_riemersma_cpu.cpp:82-92). */
int orc_tree_lookup(const OrcTree* t, const double target[4]);

/* ---- engines ---------------------------------------------------------- */
/* RiemersmaWalkCpu (rd_riemersma_cpu.cpp:359-417), palette = tree colormap.
 * rgba is dithered in place. */
void orc_walk_seq(const OrcTree* t, float* rgba, int64_t w, int64_t h,
                  double diffusion, int use_cache);
/* Same walk, but also records the palette index chosen at every visit in
 * visit order (w*h entries).  Used to check index outputs. */
void orc_walk_seq_index(const OrcTree* t, float* rgba, int64_t w, int64_t h,
                        double diffusion, int use_cache, int32_t* index_out);
/* Memo misses (full octree searches) of one cached walk (rgba dithered in place). */
int64_t orc_walk_seq_misses(const OrcTree* t, float* rgba, int64_t w, int64_t h,
                            double diffusion);
/* RiemersmaBlocksCpu unfused path (rd_riemersma_cpu.cpp:477-805).  batch is
 * frames*w*h*4 floats dithered in place; index_out (frames*n bytes, n = block
 * curve length) receives the chosen index per curve position when non-null.
 * Returns 0, or -1 for the reference's refusals (count>256, frames<1, B<16). */
int orc_walk_blocks(const OrcTree* t, float* batch, int64_t w, int64_t h,
                    int frames, int block, double diffusion,
                    uint8_t* index_out);
/* rgba64le output of the fused scatter (rd_riemersma_cpu.cpp:739-752, 794-801):
 * frames*w*h*4 uint16 from the same walk as orc_walk_blocks.  batch is read
 * only. */
int orc_walk_blocks_u16(const OrcTree* t, const float* batch, int64_t w,
                        int64_t h, int frames, int block, double diffusion,
                        uint16_t* out_u16);

/* Planar 8-bit YUV frames (fmt 7: 4:4:4, Y/Cb/Cr planes of w*h; fmt 8: 4:2:0,
 * Y then Cb, Cr at (w/2)*(h/2), even w and h) widened to RgbaF through
 * swscale's matrix (include/rd_types.h:75-97), alpha 65535; 4:2:0 chroma
 * upsampled as src/rd_blocks_cuda.cu:651-740.  Returns 0 or -1. */
int orc_yuv_to_rgbaf(const uint8_t* planes, int64_t w, int64_t h, int frames, int fmt,
                     float* out);
/* Planar 4:4:4 output of the fused scatter (rd_riemersma_cpu.cpp:711-730,
 * 776-792): frames*3*w*h bytes from the same walk as orc_walk_blocks. */
int orc_walk_blocks_yuv444(const OrcTree* t, const float* batch, int64_t w, int64_t h,
                           int frames, int block, double diffusion, uint8_t* out);

/* The threshold-screen plugin dithers in place on RgbaF frames (screen 0
 * bayer-ordered, 1 void-and-cluster, 2 void-and-cluster-fast, 3 clustered-dot;
 * src/rd_dithers_ordered.cc, src/rd_dithers.cc:294-358); ent is count x 4
 * doubles.  Returns 0 or -1. */
int orc_dither_screen(const double* ent, int count, float* rgba, int64_t w, int64_t h,
                      int frames, int screen, double diffusion);

/* The error-diffusion plugin dithers in place on RgbaF frames (kind 0
 * floyd-steinberg, 1 atkinson, 2 jarvis; src/rd_dithers.cc:74-220); ent is
 * count x 4 doubles.  Returns 0 or -1. */
int orc_dither_diffuse(const double* ent, int count, float* rgba, int64_t w, int64_t h,
                       int frames, int kind, double diffusion);

/* Inverse filter g[0..taps] (src/rd_approx_cuda.cu:73-88). */
void orc_inverse_filter(int taps, double diffusion, double* g);
/* The Jacobi engine restated on the CPU (src/rd_approx_cuda.cu:90-384):
 * gather, InitIndex, `iterations` sweeps, owner-filtered scatter into rgba.
 * fp64 selects the double field; use_tree the octree lookup.  q_out (n bytes,
 * n = block-curve length) receives the final index per curve position when
 * non-null.  Returns 0 or -1 (count>256, taps outside 1..4096). */
int orc_jacobi(const OrcTree* t, float* rgba, int64_t w, int64_t h,
               int iterations, int taps, int fp64, int use_tree,
               double diffusion, uint8_t* q_out);
/* InitIndexKernel alone (src/rd_approx_cuda.cu:127-148): q[i] = the entry
 * nearest the curve-order input cx[i] (n x 4 floats, or doubles with fp64). */
int orc_jacobi_init(const OrcTree* t, const void* cx, int64_t n, int fp64, int use_tree,
                    uint8_t* q);
/* One JacobiStepKernel sweep (src/rd_approx_cuda.cu:156-228) from a given
 * q_in into q_out (distinct arrays: the reference ping-pongs, :356-368).
 * t_out (n x 4 doubles, may be null) receives each target before clamping.
 * mutant != 0 plants a known mistake for the tests (1 in-place update, 2 last
 * tap dropped, 3 residual sign flipped, 4 residual one position late). */
int orc_jacobi_sweep(const OrcTree* t, const void* cx, int64_t n, const uint8_t* q_in,
                     uint8_t* q_out, double* t_out, int taps, int fp64, int use_tree,
                     double diffusion, int mutant);

/* The palette montage of VideoBuildPalette (src/rd_video.cpp:2373-2470,
 * :2636-2660): `sampled` frames of the decoder's bytes (fmt 1 rgb24, 2 rgba,
 * 3 rgb48, 4 rgba64le) placed in a grid `cols` wide (0: ceil(sqrt(want))),
 * whole (tile <= 0) or tile x tile nearest-neighbour cells, cells past the
 * last sample repeating it.  out (mw*mh RgbaF) may be null to query the size.
 * Returns 0 or -1. */
int orc_video_montage(const void* const* frames, int sampled, int fmt, int64_t w, int64_t h,
                      int want, int tile, int cols, float* out, int64_t* mw, int64_t* mh);

/* Thread-pool CPU baseline: walks `n` independent images (each with its own
 * tree) with `threads` std::threads, largest first.  Returns 0. */
int orc_walk_seq_many(OrcTree* const* trees, float* const* images,
                      const int64_t* ws, const int64_t* hs, int n,
                      double diffusion, int threads);

#ifdef __cplusplus
}
#endif

#endif  /* STIPPLE_ORACLE_H_ */
