/*
 * pt_kernels.hip -- CDNA4 (gfx950) kernels of the PaintTween synthesis hot
 * path.  Each kernel cites the reference kernel it implements
 * (src/PaintTweenCUDA/texSynth_kernel.cu unless noted).
 *
 * Numerics: compiled with -ffp-contract=off; every fused multiply-add is an
 * explicit fmaf, and all other arithmetic is the same sequence of IEEE
 * single-precision operations as the host CPU build of the algorithm, so the
 * offset field, canvas and residual match it bit for bit.  Texture fetches
 * are done in software: clamp addressing, point = floor(x + 0.5), bilinear
 * with FP32 weights from the texel-space coordinate.
 */
#include "pt_internal.h"
#include "pt_math.h"

#include <string.h>

#define PT_BX 16
#define PT_BY 16
/* Bandwidth-bound per-pixel kernels (sync/canvas, upsample, advection,
 * merges, random init) run 64 x 4 blocks: a wave covers 64 consecutive
 * pixels of one row, so its float4 / float2 accesses are whole cache lines. */
#define PT_RX 64
#define PT_RY 4
/* Every workgroup is 256 lanes in one dimension; a kernel that works on a
 * BX-wide 2-D block of pixels takes lane (id % BX, id / BX), the layout a
 * BX x (256 / BX) block would have (x fastest), so each wave covers the same
 * pixels.  One-dimensional workgroups let the persistent batch kernel run
 * every operator's tile body. */
#define PT_LX(BX) ((int)threadIdx.x % (BX))
#define PT_LY(BX) ((int)threadIdx.x / (BX))

/* The LDS of a tile body: in a kernel of its own, static __shared__ arrays;
 * in the persistent batch kernel, which runs many tile bodies one after
 * another in one workgroup, the members of the body's LDS struct (Lds) in
 * one arena that all of them share (template flag AR).  The kernels of their
 * own keep separate arrays: with one struct the compiler kept the exemplar's
 * buffer resource in VGPRs in the windowed propagate and re-read it into
 * SGPRs around every gather (15% slower). */
#define PT_LDS_ARRAY(TYPE, NAME, N)                                                                       \
    TYPE(*NAME##_p)[N];                                                                                   \
    if constexpr (AR) {                                                                                   \
        NAME##_p = &reinterpret_cast<Lds *>(arena)->NAME;                                                 \
    } else {                                                                                              \
        __shared__ TYPE NAME##_s[N];                                                                      \
        NAME##_p = &NAME##_s;                                                                             \
    }                                                                                                     \
    TYPE(&NAME)[N] = *NAME##_p;
#define PT_LDS_VAR(TYPE, NAME)                                                                            \
    TYPE *NAME##_p;                                                                                       \
    if constexpr (AR) {                                                                                   \
        NAME##_p = &reinterpret_cast<Lds *>(arena)->NAME;                                                 \
    } else {                                                                                              \
        __shared__ TYPE NAME##_s;                                                                         \
        NAME##_p = &NAME##_s;                                                                             \
    }                                                                                                     \
    TYPE &NAME = *NAME##_p;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ------------------------------------------------------------------ */
/* software texture fetches                                            */
/* ------------------------------------------------------------------ */
struct Bilin {
    int i00, i10, i01, i11;
    float w00, w10, w01, w11;
};

__device__ __forceinline__ Bilin bilin(int pitch, int w, int h, float x, float y)
{
    Bilin q;
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    q.w00 = (1.0f - a) * (1.0f - b);
    q.w10 = a * (1.0f - b);
    q.w01 = (1.0f - a) * b;
    q.w11 = a * b;
    q.i00 = ya * pitch + xa;
    q.i10 = ya * pitch + xb;
    q.i01 = yb * pitch + xa;
    q.i11 = yb * pitch + xb;
    return q;
}

__device__ __forceinline__ float mix1(const Bilin &q, float t00, float t10, float t01, float t11)
{
    return fmaf(q.w11, t11, fmaf(q.w01, t01, fmaf(q.w10, t10, q.w00 * t00)));
}

__device__ __forceinline__ float4 mix4(const Bilin &q, float4 t00, float4 t10, float4 t01, float4 t11)
{
    float4 r;
    r.x = mix1(q, t00.x, t10.x, t01.x, t11.x);
    r.y = mix1(q, t00.y, t10.y, t01.y, t11.y);
    r.z = mix1(q, t00.z, t10.z, t01.z, t11.z);
    r.w = mix1(q, t00.w, t10.w, t01.w, t11.w);
    return r;
}

__device__ __forceinline__ float4 tex4_lin(const float4 *t, int pitch, int w, int h, float x, float y)
{
    Bilin q = bilin(pitch, w, h, x, y);
    return mix4(q, t[q.i00], t[q.i10], t[q.i01], t[q.i11]);
}

__device__ __forceinline__ float2 tex2_lin(const float2 *t, int pitch, int w, int h, float x, float y)
{
    Bilin q = bilin(pitch, w, h, x, y);
    float2 t00 = t[q.i00], t10 = t[q.i10], t01 = t[q.i01], t11 = t[q.i11];
    return make_float2(mix1(q, t00.x, t10.x, t01.x, t11.x), mix1(q, t00.y, t10.y, t01.y, t11.y));
}

__device__ __forceinline__ float tex1_lin(const float *t, int pitch, int w, int h, float x, float y)
{
    Bilin q = bilin(pitch, w, h, x, y);
    return mix1(q, t[q.i00], t[q.i10], t[q.i01], t[q.i11]);
}

__device__ __forceinline__ int point_index(int w, int h, float x, float y)
{
    int xi = clampi((int)floorf(x + 0.5f), 0, w - 1);
    int yi = clampi((int)floorf(y + 0.5f), 0, h - 1);
    return yi * w + xi;
}

/* Luminance shift (PatchXF::luminanceShift, types.h:98-114): a per-pixel
 * offset y added to the R, G and B of the exemplar output in the canvas
 * (toColor4(c, lum), texSynth_util.h:207-209; its alpha part is always 0
 * here).  The plane is kept only while A.lum_live. */
__device__ __forceinline__ float4 lum_shifted(const PtArgs &A, float4 c, int o)
{
    if (!A.lum_live) return c;
    float l = A.base.lum[o];
    return make_float4(c.x + l, c.y + l, c.z + l, c.w + 0.0f);
}

/* The working offset's shift: the base offset's when the pixel keeps it,
 * else 0 (every new candidate carries Luminance()). */
__device__ __forceinline__ void lum_set(const PtArgs &A, int o, bool keep)
{
    if (A.lum_live) A.work.lum[o] = keep ? A.base.lum[o] : 0.f;
}

/* First texel of exemplar style layer l: a layered texture clamps the layer
 * index into [0, nl - 1].  A literal 0 folds away. */
__device__ __forceinline__ int lay(const PtArgs &A, int l)
{
    return (l <= 0 ? 0 : (l < A.nl ? l : A.nl - 1)) * (A.ew * A.eh);
}

__device__ __forceinline__ float4 tex_ex_out(const PtArgs &A, float x, float y, int l = 0)
{
    Bilin q = bilin(A.ew, A.ew, A.eh, x, y);
    const float4 *e = A.ex + 2 * lay(A, l);
    return mix4(q, e[2 * q.i00 + 1], e[2 * q.i10 + 1], e[2 * q.i01 + 1], e[2 * q.i11 + 1]);
}

__device__ __forceinline__ float tex_ex_out_alpha(const PtArgs &A, float x, float y, int l = 0)
{
    Bilin q = bilin(A.ew, A.ew, A.eh, x, y);
    const float4 *e = A.ex + 2 * lay(A, l);
    return mix1(q, e[2 * q.i00 + 1].w, e[2 * q.i10 + 1].w, e[2 * q.i01 + 1].w, e[2 * q.i11 + 1].w);
}

__device__ __forceinline__ float2 scale_at(const PtArgs &A, float x, float y)
{
    return tex2_lin(A.scale, A.scw, A.scw, A.sch, x, y);
}

/* ------------------------------------------------------------------ */
/* texSynth_util.h helpers                                             */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ bool in_bounds(int x, int y, int w, int h, int border)
{
    return x >= border && y >= border && x < w - border && y < h - border;
}

/* clampToExemplar, float version (texSynth_util.h:50-57) */
__device__ __forceinline__ void clamp_ex(float &x, float &y, int w, int h)
{
    if (x < 0) x = 0;
    if (y < 0) y = 0;
    if (x >= (float)w) x = (float)(w - 1);
    if (y >= (float)h) y = (float)(h - 1);
}

__device__ __forceinline__ float ssd_alpha(float4 a, float4 b, float alpha)
{
    float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
    float s = d0 * d0;
    s = fmaf(d1, d1, s);
    s = fmaf(d2, d2, s);
    s = fmaf(alpha * d3, d3, s);
    return s;
}

/* RGBtoLAB of a colour (texSynth_util.h:220-249), alpha unchanged; the
 * fast powers give the plain conversion's floats bit for bit */
__device__ __noinline__ float4 lab4(float4 c)
{
    float l, a, b;
    pt_rgb_to_lab_fast(c.x, c.y, c.z, &l, &a, &b);
    return make_float4(l, a, b, c.w);
}

__device__ __forceinline__ float ssd4(float4 a, float4 b)
{
    float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
    float s = d0 * d0;
    s = fmaf(d1, d1, s);
    s = fmaf(d2, d2, s);
    s = fmaf(d3, d3, s);
    return s;
}

__device__ __forceinline__ float4 sub4(float4 a, float4 b)
{
    return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}

/* screenToExemplarBasis (texSynth_util.h:95-104) */
__device__ __forceinline__ void basis(float angle, float su, float sv, float &b0x, float &b0y, float &b1x,
                                      float &b1y)
{
    float s, c;
    pt_sincosf(angle, &s, &c);
    b1x = -s * sv;
    b1y = c * sv;
    b0x = c * su;
    b0y = s * su;
}

/* calculateNewExemplarCoord (texSynth_util.h:106-118) */
__device__ __forceinline__ void new_coord(float i, float j, float ex, float ey, float b0x, float b0y, float b1x,
                                          float b1y, float &ox, float &oy)
{
    float cx = i * b0x + j * b1x;
    float cy = i * b0y + j * b1y;
    ox = ex + cx;
    oy = ey + cy;
}

/* screenToExemplarAngle (texSynth_kernel.cu:425-436), style layer l */
__device__ __forceinline__ float screen_to_ex_angle(const PtArgs &A, float x, float y, float ex, float ey, int l = 0)
{
    if (A.orient_en) {
        float theta =
            A.ex_orient[lay(A, l) + point_index(A.ew, A.eh, ex, ey)] - A.orient[point_index(A.ow, A.oh, x, y)];
        if ((double)theta > 1.57079632679489661923)
            return (float)((double)theta - 3.14159265358979323846);
        if ((double)theta < -1.57079632679489661923)
            return (float)(3.14159265358979323846 + (double)theta);
        return theta;
    }
    return 0.0f;
}

/* isImagePixelCulled (texSynth_kernel.cu:176-186), integer pixel */
__device__ __forceinline__ bool image_culled(const PtArgs &A, int x, int y)
{
    return A.dt[A.cur][y * A.W + x] > ((float)A.cull - 0.0f);
}

/* isExemplarPixelCulled (texSynth_kernel.cu:200-205): its parameters are
 * int, so the candidate position is truncated toward zero and the texel at
 * that integer position read exactly (clamped). */
__device__ __forceinline__ bool exemplar_culled(const PtArgs &A, float x, float y, int l = 0)
{
    float d = A.ex_dt[lay(A, l) + clampi((int)y, 0, A.eh - 1) * A.ew + clampi((int)x, 0, A.ew - 1)];
    return (double)d > (double)A.cull * 1.25;
}

/* hysteresisTerm (texSynth_kernel.cu:506-535) */
__device__ float hysteresis_term(const PtArgs &A, int x, int y)
{
    float maxD = A.max_dist / (float)(1 << A.level);
    const float2 *v = (A.direction == PT_FORWARD) ? A.velb[A.cur] : A.velf[A.cur];
    float2 d00 = tex2_lin(v, A.W, A.W, A.H, (float)x, (float)y);
    float2 d10 = tex2_lin(v, A.W, A.W, A.H, (float)x + 1.f, (float)y);
    float2 d01 = tex2_lin(v, A.W, A.W, A.H, (float)x, (float)y + 1.f);
    float2 d11 = tex2_lin(v, A.W, A.W, A.H, (float)x + 1.f, (float)y + 1.f);
    float maxX = fmaxf(d00.x, fmaxf(d01.x, fmaxf(d10.x, d11.x)));
    float maxY = fmaxf(d00.y, fmaxf(d01.y, fmaxf(d10.y, d11.y)));
    float minX = fminf(d00.x, fminf(d01.x, fminf(d10.x, d11.x)));
    float minY = fminf(d00.y, fminf(d01.y, fminf(d10.y, d11.y)));
    if ((maxX - minX) < maxD && (maxY - minY) < maxD)
        return A.hyst_w;
    return 0.f;
}

/* Index of tap (i, j) of pixel (x, y) in a time table: tap-major planes of
 * the level's pixels, two float4 (exemplar base, output) per entry. */
__device__ __forceinline__ int time_table_index(const PtArgs &A, int x, int y, int i, int j)
{
    const int r = A.r, t = (i + r) * (2 * r + 1) + (j + r);
    return (t * (A.lw * A.lh) + y * A.lw + x) * 2;
}

/* The toroidal wrap of a tap coordinate (texSynth_kernel.cu:590-599): once
 * down when t >= h, else once up when t < 0.  Written as selects: when
 * t >= h, t - h is >= 0 exactly (rounding keeps its sign), so the second test
 * fires only when the first did not, as in the reference's if / else if; a
 * NaN fails both.  Branch-free, so a wave does not split per tap. */
__device__ __forceinline__ float wrap_tap(float t, float h)
{
    float a = t >= h ? t - h : t;
    return a < 0.f ? a + h : a;
}

/* ------------------------------------------------------------------ */
/* patch distance                                                      */
/* ------------------------------------------------------------------ */

/* The workgroup's target colour and canvas, staged in LDS with an r-pixel
 * halo: tile row pitch tw = PT_BX + 2r; the lane's pixel sits at
 * (lx + r, ly + r). */
struct Tile {
    const float4 *col;
    const float4 *can;
    int tw, lx, ly;
    /* exemplar window: ww x ww texels from (wx, wy), interleaved base/output
     * in LDS; exl: every bilinear fetch of the lane's current candidate
     * lies inside it (read only by the XL kernel variants) */
    const float4 *exw;
    int wx, wy, ww;
    bool exl;
    /* advected set sdir (-1: none) of the tile's pixels and halo, staged in
     * LDS with the tile's pitch: validity, previous-frame colour and
     * exemplar output (read only by the TS kernel variants) */
    const float4 *ab, *ao;
    const unsigned char *av;
    int sdir;
};

typedef float pt_f4v __attribute__((ext_vector_type(4)));

/* 16-byte load through a buffer resource at a 32-bit byte offset. */
__device__ __forceinline__ float4 ld4(__amdgpu_buffer_rsrc_t r, int off)
{
    pt_f4v v = __builtin_bit_cast(pt_f4v, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    return make_float4(v.x, v.y, v.z, v.w);
}

/* Buffer resource over the interleaved exemplar texture (32 B per texel),
 * every style layer. */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ex_rsrc(const PtArgs &A)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)A.ex, (short)0, (A.nl > 1 ? A.nl : 1) * A.ew * A.eh * 32,
                                             0x00020000);
}

/* Bilinear exemplar base and output at (x, y), clamp addressing; the same
 * weights and fused order as tex4_lin.  lb: first texel of the style layer. */
__device__ __forceinline__ void ex_pair(__amdgpu_buffer_rsrc_t er, int ew, int eh, float x, float y, float4 &b,
                                        float4 &o, int lb = 0)
{
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, c = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int xa = clampi(x0, 0, ew - 1), xb = clampi(x0 + 1, 0, ew - 1);
    int ya = clampi(y0, 0, eh - 1), yb = clampi(y0 + 1, 0, eh - 1);
    Bilin q;
    q.w00 = (1.0f - a) * (1.0f - c);
    q.w10 = a * (1.0f - c);
    q.w01 = (1.0f - a) * c;
    q.w11 = a * c;
    int ra = lb + ya * ew, rb = lb + yb * ew;
    int o00 = (ra + xa) << 5, o10 = (ra + xb) << 5, o01 = (rb + xa) << 5, o11 = (rb + xb) << 5;
    b = mix4(q, ld4(er, o00), ld4(er, o10), ld4(er, o01), ld4(er, o11));
    o = mix4(q, ld4(er, o00 + 16), ld4(er, o10 + 16), ld4(er, o01 + 16), ld4(er, o11 + 16));
}

/* ex_pair from the workgroup's LDS window when the lane's candidate lies
 * inside it (same texels, weights and order as the global fetch). */
template <bool XL>
__device__ __forceinline__ void ex_pair_t(const Tile &T, __amdgpu_buffer_rsrc_t er, int ew, int eh, float x, float y,
                                          float4 &b, float4 &o, int lb = 0)
{
    if constexpr (XL) {
        /* one set of weights and one mix for both sources, so that a wave
         * whose lanes read from both pays the arithmetic once */
        float fx = floorf(x), fy = floorf(y);
        float a = x - fx, c = y - fy;
        Bilin q;
        q.w00 = (1.0f - a) * (1.0f - c);
        q.w10 = a * (1.0f - c);
        q.w01 = (1.0f - a) * c;
        q.w11 = a * c;
        int x0 = (int)fx, y0 = (int)fy;
        float4 b00, b10, b01, b11, o00, o10, o01, o11;
        if (T.exl) {
            int i00 = ((y0 - T.wy) * T.ww + (x0 - T.wx)) * 2;
            int i01 = i00 + 2 * T.ww;
            const float4 *e = T.exw;
            b00 = e[i00];
            o00 = e[i00 + 1];
            b10 = e[i00 + 2];
            o10 = e[i00 + 3];
            b01 = e[i01];
            o01 = e[i01 + 1];
            b11 = e[i01 + 2];
            o11 = e[i01 + 3];
        } else {
            int xa = clampi(x0, 0, ew - 1), xb = clampi(x0 + 1, 0, ew - 1);
            int ya = clampi(y0, 0, eh - 1), yb = clampi(y0 + 1, 0, eh - 1);
            int ra = lb + ya * ew, rb = lb + yb * ew;
            int f00 = (ra + xa) << 5, f10 = (ra + xb) << 5, f01 = (rb + xa) << 5, f11 = (rb + xb) << 5;
            b00 = ld4(er, f00);
            b10 = ld4(er, f10);
            b01 = ld4(er, f01);
            b11 = ld4(er, f11);
            o00 = ld4(er, f00 + 16);
            o10 = ld4(er, f10 + 16);
            o01 = ld4(er, f01 + 16);
            o11 = ld4(er, f11 + 16);
        }
        b = mix4(q, b00, b10, b01, b11);
        o = mix4(q, o00, o10, o01, o11);
    } else {
        ex_pair(er, ew, eh, x, y, b, o, lb);
    }
}

/* Whether every tap of a candidate centred at (cx, cy) with basis (b0, b1)
 * reads its bilinear texels inside the window at (wx, wy) of side ww,
 * without wrap-around (taps in [0, hw) x [0, hh)) and without edge clamping
 * (texels inside the exemplar).  Taps lie within r(|b0|+|b1|) of the centre
 * per axis; the margin covers their rounding.  False for NaN. */
__device__ __forceinline__ bool in_window(const PtArgs &A, int r, float cx, float cy, float b0x, float b0y, float b1x,
                                          float b1y, int wx, int wy, int ww)
{
    float ex = (float)r * (fabsf(b0x) + fabsf(b1x)) + 0.01f;
    float ey = (float)r * (fabsf(b0y) + fabsf(b1y)) + 0.01f;
    float lx = cx - ex, hx = cx + ex, ly = cy - ey, hy = cy + ey;
    return lx >= (float)wx && hx < (float)(wx + ww - 1) && ly >= (float)wy && hy < (float)(wy + ww - 1) &&
           lx >= 0.f && ly >= 0.f && hx < (float)A.hw && hy < (float)A.hh && hx < (float)(A.ew - 1) &&
           hy < (float)(A.eh - 1);
}

/* Stages the exemplar window of side ww anchored on the offset (ax, ay):
 * its origin puts the anchor near the centre, clamped inside the exemplar
 * (the caller guarantees ew, eh >= ww). */
__device__ __forceinline__ void stage_window(const PtArgs &A, float ax, float ay, int ww, float4 *sw, int &wx, int &wy,
                                             int tid, int nthreads)
{
    wx = clampi((int)floorf(ax) - ww / 2, 0, A.ew - ww);
    wy = clampi((int)floorf(ay) - ww / 2, 0, A.eh - ww);
    for (int k = tid; k < ww * ww * 2; k += nthreads) {
        int t = k >> 1;
        sw[k] = A.ex[2 * ((wy + t / ww) * A.ew + wx + t % ww) + (k & 1)];
    }
}

/* timeDerivativeTerms (texSynth_kernel.cu:438-469); EXT: the kernel
 * variant with the extended terms, where the LAB colour space converts the
 * previous frame's colours and exemplar samples (:454-459).  Time-table
 * entries are stored converted when that colour space is on (only the
 * extended variants run then); PREV_LAB: the caller's previous-frame colours
 * are converted already. */
template <bool EXT = false, bool PREV_LAB = false>
__device__ float time_terms_v(const PtArgs &A, __amdgpu_buffer_rsrc_t er, int k, float4 prev, float pb0x, float pb0y,
                              float pb1x, float pb1y, float4 base, float4 tb, float4 canvas, float4 to,
                              float4 prev_base, float4 prev_out, int x, int y, int i, int j, int plb = 0)
{
    float4 peb, peo;
    if (A.tt[k]) {
        /* the exemplar samples at the advected offset's tap, gathered once
         * per level by k_time_table with this same arithmetic */
        const float4 *e = A.tt[k] + time_table_index(A, x, y, i, j);
        peb = e[0];
        peo = e[1];
    } else {
        float ptx, pty;
        new_coord((float)i, (float)j, prev.x, prev.y, pb0x, pb0y, pb1x, pb1y, ptx, pty);
        ex_pair(er, A.ew, A.eh, ptx, pty, peb, peo, plb);
        if (EXT && A.lab) {
            peb = lab4(peb);
            peo = lab4(peo);
        }
    }
    if (EXT && !PREV_LAB && A.lab) {
        prev_base = lab4(prev_base);
        prev_out = lab4(prev_out);
    }
    float r = ssd4(sub4(base, prev_base), sub4(tb, peb)) * A.tdi;
    r = fmaf(ssd4(sub4(canvas, prev_out), sub4(to, peo)), A.tdo, r);
    r = fmaf(ssd4(prev_out, to), A.tcw, r);
    return r;
}

/* time_terms_v with the advected colours of tap pixel oo read from set k */
template <bool EXT = false>
__device__ __forceinline__ float time_terms_b(const PtArgs &A, __amdgpu_buffer_rsrc_t er, int k, float4 prev,
                                              float pb0x, float pb0y, float pb1x, float pb1y, float4 base, float4 tb,
                                              float4 canvas, float4 to, int oo, int x, int y, int i, int j,
                                              int plb = 0)
{
    const PtAdv &d = A.adv[k];
    return time_terms_v<EXT>(A, er, k, prev, pb0x, pb0y, pb1x, pb1y, base, tb, canvas, to, d.base[oo], d.out[oo], x, y,
                             i, j, plb);
}

/* pt_coherence_term_k as a called function: the tap loops that evaluate it
 * keep their registers for the gathers */
__device__ __noinline__ float coherence_term(float cx, float cy, float tcx, float tcy, float tx, float ty, int i, int j,
                                             int level, float inv_ratio, float ratio, float z, double lim)
{
    return pt_coherence_term_k(cx, cy, tcx, tcy, tx, ty, i, j, level, inv_ratio, ratio, z, lim);
}

/* The coherence term of tap (i, j) (texSynth_kernel.cu:642-646): against
 * the current offset of pixel (x+i, y+j); (tcx, tcy) is the candidate's
 * centre and (tx, ty) the wrapped tap position. */
__device__ __forceinline__ float coherence_add(const PtArgs &A, float res, int x, int y, int i, int j, float tcx,
                                               float tcy, float tx, float ty)
{
    if (A.coh_w > 0.f) {
        float4 n = A.base.xyth[(y + j) * A.W + (x + i)];
        res = res + coherence_term(n.x, n.y, tcx, tcy, tx, ty, i, j, A.level, A.coh_inv_ratio, A.coh_ratio, A.coh_z,
                                   A.coh_lim) *
                        A.coh_w;
    }
    return res;
}

/* Radius codes of the kernel templates: RT = 2, 3, 4 a patch radius fixed
 * at compile time, 0 a run-time radius; RT < 0 the variants with the
 * extended patch-distance terms (coherence, LAB colour space, object IDs):
 * -1 at a run-time radius (the per-pixel kernels, style layers included),
 * -2 and -3 at radius 2 and 3 (the unit-parallel kernels, without style
 * layers).  The default kernels (RT >= 0) leave those terms out and keep
 * their registers. */
template <int RT> struct Rt {
    static constexpr int R = RT > 0 ? RT : (RT < -1 ? -RT : 0);
    static constexpr bool EXT = RT < 0;
};
#define PT_RAD(RT, r) (Rt<RT>::R > 0 ? Rt<RT>::R : (r))

/*
 *
 * calculateResidual (texSynth_kernel.cu:537-659): one patch distance of
 * candidate (t, tsc) at pixel (x, y).  Taps run i over x outer, j over y
 * inner and are accumulated sequentially per lane in that order.  TIME
 * selects the variant with the temporal terms (an advected offset is valid at
 * the centre); both variants perform the same operations when those terms are
 * inactive.  Target and canvas come from the LDS tile; i*b0 is formed once
 * per row, exactly the product the reference forms per tap.
 */
template <int RT, bool TIME, bool XL = false>
__device__ float calc_residual_impl(const PtArgs &A, const Tile &T, int x, int y, float4 t, float2 tsc, float4 pf,
                                    float4 pb, bool cpv, bool cnv, int tl = 0, int pfl = 0, int pbl = 0)
{
    const int r = PT_RAD(RT, A.r);
    /* style layers of the candidate and of the advected offsets (RT < 0) */
    const int tlb = RT < 0 ? lay(A, tl) : 0;
    const int pflb = RT < 0 ? lay(A, pfl) : 0, pblb = RT < 0 ? lay(A, pbl) : 0;
    __amdgpu_buffer_rsrc_t er = ex_rsrc(A);
    const int ew = A.ew, eh = A.eh;
    const float hw = (float)A.hw, hh = (float)A.hh;
    int o = y * A.W + x;
    bool any_time_w = A.tdi > 0.f || A.tdo > 0.f || A.tcw > 0.f;
    bool fwd_dir = A.direction == PT_FORWARD || A.direction == PT_BIDIRECTIONAL;
    bool bwd_dir = A.direction == PT_BACKWARD || A.direction == PT_BIDIRECTIONAL;
    float fb0x = 0, fb0y = 0, fb1x = 0, fb1y = 0, bb0x = 0, bb0y = 0, bb1x = 0, bb1y = 0;
    float sum_w = 0.f, sum_r = 0.f;
    if (TIME) {
        if (cpv) {
            float2 ps = A.adv[0].sc[o];
            float ang = screen_to_ex_angle(A, (float)x, (float)y, pf.x, pf.y, RT < 0 ? pfl : 0);
            basis(ang, A.su * ps.x, A.sv * ps.y, fb0x, fb0y, fb1x, fb1y);
        }
        if (cnv) {
            float2 ps = A.adv[1].sc[o];
            float ang = screen_to_ex_angle(A, (float)x, (float)y, pb.x, pb.y, RT < 0 ? pbl : 0);
            basis(ang, A.su * ps.x, A.sv * ps.y, bb0x, bb0y, bb1x, bb1y);
        }
        if (cpv)
            sum_r -= hysteresis_term(A, x, y);
    }
    float b0x, b0y, b1x, b1y;
    basis(t.z, A.su * tsc.x, A.sv * tsc.y, b0x, b0y, b1x, b1y);
#pragma unroll 1
    for (int i = -r; i <= r; i++) {
        if (x + i < 0 || x + i >= A.lw)
            continue;
        const float ibx = (float)i * b0x, iby = (float)i * b0y;
        const float4 *tcol = T.col + (T.ly + r) * T.tw + (T.lx + r + i);
        const float4 *tcan = T.can + (T.ly + r) * T.tw + (T.lx + r + i);
#pragma unroll
        for (int j = -r; j <= r; j++) {
            if (y + j < 0 || y + j >= A.lh)
                continue;
            /* staged converted to LAB when the colour space is on */
            float4 base = tcol[j * T.tw];
            float4 canvas = tcan[j * T.tw];
            float tx = t.x + (ibx + (float)j * b1x);
            float ty = t.y + (iby + (float)j * b1y);
            tx = wrap_tap(tx, hw);
            ty = wrap_tap(ty, hh);
            float4 tb, to;
            ex_pair_t<XL>(T, er, ew, eh, tx, ty, tb, to, tlb);
            if (RT < 0 && A.lab) {
                tb = lab4(tb);
                to = lab4(to);
            }
            float res = ssd_alpha(base, tb, A.alpha_w) * A.in_w;
            if (TIME && any_time_w) {
                int oo = (y + j) * A.W + (x + i);
                if (fwd_dir && cpv && A.adv[0].xyth[oo].w > 0)
                    res = res + time_terms_b<(RT < 0)>(A, er, 0, pf, fb0x, fb0y, fb1x, fb1y, base, tb, canvas, to, oo,
                                                       x, y, i, j, pflb);
                if (bwd_dir && cnv && A.adv[1].xyth[oo].w > 0)
                    res = res + time_terms_b<(RT < 0)>(A, er, 1, pb, bb0x, bb0y, bb1x, bb1y, base, tb, canvas, to, oo,
                                                       x, y, i, j, pblb);
            }
            if (i == 0 && j == 0) {
                if (A.dt_w > 0.f) {
                    float de = tex1_lin(A.ex_dt + tlb, A.ew, A.ew, A.eh, tx, ty);
                    de /= tsc.y;
                    float di = A.dt[A.cur][o];
                    float diff = fabsf(de - di);
                    diff = fminf(fmaxf(diff, 0.0f), 10.0f);
                    res = res + diff * A.dt_w;
                }
                /* objectIdTerm (texSynth_kernel.cu:497-504, 637-638): point
                 * fetches, the image ID at the pixel, the style's at the tap */
                if (RT < 0 && A.obj_w > 0.f) {
                    int iid = A.img_id[point_index(A.idw, A.idh, (float)x, (float)y)];
                    int sid = A.ex_id[tlb + point_index(A.ew, A.eh, tx, ty)];
                    if (iid != sid) res = res + A.obj_w;
                }
            } else {
                res = fmaf(ssd_alpha(canvas, to, A.alpha_w), A.canvas_w, res);
                if (RT < 0) res = coherence_add(A, res, x, y, i, j, t.x, t.y, tx, ty);
            }
            float wgt = A.gauss[i + r] * A.gauss[j + r];
            sum_r = fmaf(res, wgt, sum_r);
            sum_w += wgt;
        }
    }
    if (sum_w < 0.000001f)
        return 1e31f;
    return sum_r / sum_w;
}

/* TK (a kernel-level switch) is false when the host knows no advected
 * offset is valid (the advected buffers were reset and not written since);
 * the per-pixel test then always takes the plain variant. */
template <int RT, bool TK, bool XL = false>
__device__ __forceinline__ float calc_residual(const PtArgs &A, const Tile &T, int x, int y, float4 t, float2 tsc,
                                               int tl = 0)
{
    if (TK) {
        int o = y * A.W + x;
        float4 pf = A.adv[0].xyth[o], pb = A.adv[1].xyth[o];
        bool cpv = pf.w > 0, cnv = pb.w > 0;
        int pfl = 0, pbl = 0;
        if (RT < 0 && A.layered) {
            pfl = A.adv[0].la[o];
            pbl = A.adv[1].la[o];
        }
        if (cpv || cnv)
            return calc_residual_impl<RT, true, XL>(A, T, x, y, t, tsc, pf, pb, cpv, cnv, tl, pfl, pbl);
    }
    float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    return calc_residual_impl<RT, false, XL>(A, T, x, y, t, tsc, z, z, false, false, tl);
}

/* test_offset (texSynth_kernel.cu:695-721); tl: the candidate's style
 * layer (RT < 0 only) */
template <int RT, bool TK>
__device__ __forceinline__ float test_offset(const PtArgs &A, const Tile &T, int x, int y, float4 t, float2 tsc,
                                             float image_alpha, int tl = 0)
{
    if (RT >= 0) tl = 0;
    if (!exemplar_culled(A, t.x, t.y, tl)) {
        if (!A.transp) {
            float ea = tex_ex_out_alpha(A, t.x, t.y, tl);
            if (ea < 0.99f && image_alpha > 0.99f)
                return 1e31f;
        }
        return calc_residual<RT, TK>(A, T, x, y, t, tsc, tl);
    }
    return 1e31f;
}

/* ------------------------------------------------------------------ */
/* kernels                                                             */
/* ------------------------------------------------------------------ */
/* The tile of this workgroup in a single-job launch.  Workgroups are dealt
 * to the 8 XCDs round-robin; with xcd_tiles XCD x takes the contiguous
 * row-major range of tiles starting at sum_{y<x} count_y, so neighbouring
 * tiles (whose candidates gather neighbouring exemplar lines) share one L2. */
__device__ __forceinline__ void xcd_tile(const PtArgs &A, int &bx, int &by)
{
    bx = blockIdx.x;
    by = blockIdx.y;
    if (!A.xcd_tiles) return;
    const int nx = gridDim.x, n = gridDim.x * gridDim.y;
    int g = blockIdx.x + blockIdx.y * nx;
    int q = n >> 3, rem = n & 7, x = g & 7, k = g >> 3;
    g = x * q + (x < rem ? x : rem) + k;
    bx = g % nx;
    by = g / nx;
}

__device__ __forceinline__ bool pixel_of_rows(const PtArgs &A, int bx, int by, int &x, int &y)
{
    x = bx * PT_RX + PT_LX(PT_RX);
    y = by * PT_RY + PT_LY(PT_RX);
    return x < A.lw && y < A.lh;
}

__device__ __forceinline__ bool pixel_of(const PtArgs &A, int bx, int by, int &x, int &y)
{
    x = bx * PT_BX + PT_LX(PT_BX);
    y = by * PT_BY + PT_LY(PT_BX);
    return x < A.lw && y < A.lh;
}

template <int RT> struct TileSize {
    static constexpr int R = PT_RAD(RT, PT_MAX_R);
    static constexpr int N = (PT_BX + 2 * R) * (PT_BY + 2 * R);
};

/* Stages the block's target colour and canvas (with an r halo, coordinates
 * clamped into the level; clamped entries are never read because taps
 * outside the level are skipped) and returns the lane's view.  With the LAB
 * colour space (RT < 0) both are staged converted: every tap reading the
 * entry converts that same colour (texSynth_kernel.cu:575-578). */
template <int RT>
__device__ __forceinline__ Tile stage_tile(const PtArgs &A, int bx, int by, float4 *scol, float4 *scan)
{
    const int r = PT_RAD(RT, A.r);
    const int tw = PT_BX + 2 * r, th = PT_BY + 2 * r;
    const int x0 = bx * PT_BX - r, y0 = by * PT_BY - r;
    const float4 *color = A.color[A.cur];
    const bool lab = RT < 0 && A.lab;
    for (int k = threadIdx.x; k < tw * th; k += PT_BX * PT_BY) {
        int tx = clampi(x0 + k % tw, 0, A.lw - 1);
        int ty = clampi(y0 + k / tw, 0, A.lh - 1);
        float4 c = color[ty * A.W + tx], v = A.canvas[ty * A.W + tx];
        if (lab) {
            c = lab4(c);
            v = lab4(v);
        }
        scol[k] = c;
        scan[k] = v;
    }
    __syncthreads();
    Tile T;
    T.col = scol;
    T.can = scan;
    T.tw = tw;
    T.lx = PT_LX(PT_BX);
    T.ly = PT_LY(PT_BX);
    return T;
}

/*
 * updateResidualCache_kernel (texSynth_kernel.cu:661-693).
 *
 * Incremental form: a pixel's residual depends on its own offset, the
 * canvas inside its patch window (each canvas value depends only on that
 * pixel's offset), and inputs that do not change between two refreshes of
 * one level.  When the residual cache was exact before the last sync and no
 * offset inside the window changed in that sync (A.changed, written by the
 * sync pass), recomputing would reproduce the stored value bit for bit, so
 * the pixel is skipped.  Pixels that need the refresh are compacted onto the
 * first lanes of the workgroup so that whole waves skip.
 */
#define PT_XW_RES 24 /* exemplar window side of the XL residual refresh */

template <int RT, bool XL> struct ResLds {
    float4 scol[TileSize<RT>::N], scan[TileSize<RT>::N];
    float4 sexw[XL ? PT_XW_RES * PT_XW_RES * 2 : 1];
    unsigned char sflag[TileSize<RT>::N];
    unsigned short slist[PT_BX * PT_BY];
    int scount, scount_b;
};

template <int RT, bool TK, bool XL, bool AR = false>
__device__ __forceinline__ void b_update_residual(const PtArgs &A, int bx, int by, void *arena = nullptr)
{
    using Lds = ResLds<RT, XL>;
    PT_LDS_ARRAY(float4, scol, TileSize<RT>::N)
    PT_LDS_ARRAY(float4, scan, TileSize<RT>::N)
    PT_LDS_ARRAY(float4, sexw, (XL ? PT_XW_RES * PT_XW_RES * 2 : 1))
    PT_LDS_ARRAY(unsigned char, sflag, TileSize<RT>::N)
    PT_LDS_ARRAY(unsigned short, slist, PT_BX * PT_BY)
    PT_LDS_VAR(int, scount)
    PT_LDS_VAR(int, scount_b)
    const int r = PT_RAD(RT, A.r);
    const bool inc = A.res_incremental != 0;
    if (inc) {
        const int tw = PT_BX + 2 * r, th = PT_BY + 2 * r;
        const int x0 = bx * PT_BX - r, y0 = by * PT_BY - r;
        for (int k = threadIdx.x; k < tw * th; k += PT_BX * PT_BY) {
            int tx = clampi(x0 + k % tw, 0, A.lw - 1);
            int ty = clampi(y0 + k / tw, 0, A.lh - 1);
            sflag[k] = A.changed[ty * A.W + tx];
        }
        if (threadIdx.x == 0) {
            scount = 0;
            scount_b = 0;
        }
    }
    int x, y;
    bool valid = pixel_of(A, bx, by, x, y);
    Tile T = stage_tile<RT>(A, bx, by, scol, scan); /* its barrier also covers sflag and scount */
    /* XL: the window is anchored on the offset of the tile's middle pixel
     * (clamped into the level) and read by the pixels whose patch lies in it */
    auto window = [&]() {
        const int tid = threadIdx.x;
        int ax = min(bx * PT_BX + PT_BX / 2, A.lw - 1), ay = min(by * PT_BY + PT_BY / 2, A.lh - 1);
        float4 an = A.base.xyth[ay * A.W + ax];
        stage_window(A, an.x, an.y, PT_XW_RES, sexw, T.wx, T.wy, tid, PT_BX * PT_BY);
        T.exw = sexw;
        T.ww = PT_XW_RES;
        __syncthreads();
    };
    auto fits = [&](float4 t, float2 tsc) {
        float b0x, b0y, b1x, b1y;
        basis(t.z, A.su * tsc.x, A.sv * tsc.y, b0x, b0y, b1x, b1y);
        return in_window(A, r, t.x, t.y, b0x, b0y, b1x, b1y, T.wx, T.wy, T.ww);
    };
    if (!inc) {
        if (XL) window();
        if (!valid) return;
        int o = y * A.W + x;
        if (image_culled(A, x, y)) {
            A.res_base[o] = 0.0f;
            return;
        }
        float4 t = A.base.xyth[o];
        float2 tsc = A.base.sc[o];
        if (XL) T.exl = fits(t, tsc);
        A.res_base[o] = calc_residual<RT, TK, XL>(A, T, x, y, t, tsc, (RT < 0 && A.layered) ? A.base.la[o] : 0);
        return;
    }
    if constexpr (XL) {
        /* the window's origin (stage_window's), for the fit test below */
        int ax = min(bx * PT_BX + PT_BX / 2, A.lw - 1), ay = min(by * PT_BY + PT_BY / 2, A.lh - 1);
        float4 an = A.base.xyth[ay * A.W + ax];
        T.wx = clampi((int)floorf(an.x) - PT_XW_RES / 2, 0, A.ew - PT_XW_RES);
        T.wy = clampi((int)floorf(an.y) - PT_XW_RES / 2, 0, A.eh - PT_XW_RES);
        T.ww = PT_XW_RES;
    }
    if (valid) {
        /* clamped halo entries repeat edge flags: conservative */
        bool need = false;
        for (int j = 0; j <= 2 * r && !need; j++)
            for (int i = 0; i <= 2 * r; i++)
                if (sflag[(PT_LY(PT_BX) + j) * T.tw + PT_LX(PT_BX) + i]) {
                    need = true;
                    break;
                }
        if (need) {
            /* XL: pixels whose patch lies in the window first, the others
             * from the end, so that a wave mostly issues one kind of gather */
            bool front = true;
            if constexpr (XL) {
                int o = y * A.W + x;
                front = !image_culled(A, x, y) && fits(A.base.xyth[o], A.base.sc[o]);
            }
            if (front) slist[atomicAdd(&scount, 1)] = (unsigned short)threadIdx.x;
            else slist[PT_BX * PT_BY - 1 - atomicAdd(&scount_b, 1)] = (unsigned short)threadIdx.x;
        }
    }
    __syncthreads();
    const int na = scount;
    const int n = XL ? na + scount_b : na;
    if (XL && n > 0) window();
    const int lane = threadIdx.x;
    if (lane >= n) return;
    const int p = slist[(!XL || lane < na) ? lane : PT_BX * PT_BY - 1 - (lane - na)];
    x = bx * PT_BX + p % PT_BX;
    y = by * PT_BY + p / PT_BX;
    T.lx = p % PT_BX;
    T.ly = p / PT_BX;
    int o = y * A.W + x;
    if (image_culled(A, x, y)) {
        A.res_base[o] = 0.0f;
        return;
    }
    float4 t = A.base.xyth[o];
    float2 tsc = A.base.sc[o];
    if (XL) T.exl = lane < na; /* the pixels listed in front fit the window */
    A.res_base[o] = calc_residual<RT, TK, XL>(A, T, x, y, t, tsc, (RT < 0 && A.layered) ? A.base.la[o] : 0);
}

template <int RT, bool TK, bool XL = false>
__global__ __launch_bounds__(PT_BX *PT_BY) void k_update_residual(PtArgs A)
{
    int bx, by;
    xcd_tile(A, bx, by);
    b_update_residual<RT, TK, XL>(A, bx, by);
}

/* offsetsToOutput (texSynth_kernel.cu:263-274) */
__device__ __forceinline__ void b_canvas(const PtArgs &A, int bx, int by)
{
    int x, y;
    if (!pixel_of_rows(A, bx, by, x, y)) return;
    int o = y * A.W + x;
    float4 n = A.base.xyth[o];
    A.canvas[o] = lum_shifted(A, tex_ex_out(A, n.x, n.y, A.layered ? A.base.la[o] : 0), o);
    if (A.changed) A.changed[o] = 0;
}

__global__ __launch_bounds__(PT_BX *PT_BY) void k_canvas(PtArgs A)
{
    b_canvas(A, blockIdx.x, blockIdx.y);
}

/* syncBuffers' copyWorkingToBase + offsetsToOutput in one pass
 * (texSynth_kernel.cu:1244-1246,1274-1275): pixels of the copy rectangle
 * (dw x dh) take working -> base; pixels of the level rectangle then get
 * the canvas of their (new) base offset. */
__device__ __forceinline__ bool bits_differ(float a, float b) { return __float_as_uint(a) != __float_as_uint(b); }

__device__ __forceinline__ void b_sync_canvas(const PtArgs &A, int bx, int by, int dw, int dh)
{
    int x = bx * PT_RX + PT_LX(PT_RX);
    int y = by * PT_RY + PT_LY(PT_RX);
    int o = y * A.W + x;
    float4 n;
    int nl = 0;
    bool have = false, ch = false;
    if (x < dw && y < dh) {
        n = A.work.xyth[o];
        float2 s = A.work.sc[o];
        if (A.changed) {
            float4 ob = A.base.xyth[o];
            float2 os = A.base.sc[o];
            ch = bits_differ(n.x, ob.x) || bits_differ(n.y, ob.y) || bits_differ(n.z, ob.z) ||
                 bits_differ(n.w, ob.w) || bits_differ(s.x, os.x) || bits_differ(s.y, os.y);
        }
        A.base.xyth[o] = n;
        A.base.sc[o] = s;
        if (A.lum_live) {
            float l = A.work.lum[o];
            ch = ch || (A.changed && bits_differ(l, A.base.lum[o]));
            A.base.lum[o] = l;
        }
        if (A.layered) {
            nl = A.work.la[o];
            ch = ch || (A.changed && nl != A.base.la[o]);
            A.base.la[o] = nl;
        }
        have = true;
    }
    if (x < A.lw && y < A.lh) {
        if (!have) {
            n = A.base.xyth[o];
            if (A.layered) nl = A.base.la[o];
        }
        A.canvas[o] = lum_shifted(A, tex_ex_out(A, n.x, n.y, nl), o);
        if (A.changed) A.changed[o] = ch ? 1 : 0;
    }
}

__global__ __launch_bounds__(PT_BX *PT_BY) void k_sync_canvas(PtArgs A, int dw, int dh)
{
    b_sync_canvas(A, blockIdx.x, blockIdx.y, dw, dh);
}

/* copyRandomKernel (texSynth_kernel.cu:2044-2059) */
__device__ __forceinline__ void b_copy_random(const PtArgs &A, int bx, int by)
{
    int x, y;
    if (!pixel_of_rows(A, bx, by, x, y)) return;
    int o = y * A.W + x;
    float4 r = A.random[point_index(A.rw, A.rh, (float)x, (float)y)];
    A.work.xyth[o] = make_float4(r.x, r.y, r.z, 0.f);
    A.work.sc[o] = make_float2(1.f, 1.f);
    lum_set(A, o, false);
    /* toPatchXF(random): the style layer is the texel's w (texSynth_util.h:212-214) */
    if (A.layered) A.work.la[o] = (int)r.w;
}

__global__ __launch_bounds__(PT_BX *PT_BY) void k_copy_random(PtArgs A) { b_copy_random(A, blockIdx.x, blockIdx.y); }

/* calculatePropagateOffset (texSynth_kernel.cu:956-982); pixelWrap
 * (cudaImageBuffer.h:81-87) wraps modulo the level-0 dimensions.  With cl
 * given, the candidate keeps the neighbour's style layer (*cl). */
__device__ __forceinline__ void propagate_offset(const PtArgs &A, int tx, int ty, float x, float y, float4 &c,
                                                 float2 &csc, int *cl = nullptr)
{
    int wx = (int)((float)tx - (float)A.W * floorf((float)tx / (float)A.W));
    int wy = (int)((float)ty - (float)A.H * floorf((float)ty / (float)A.H));
    float4 t = A.base.xyth[wy * A.W + wx];
    int l = 0;
    if (cl) *cl = l = A.base.la[wy * A.W + wx];
    float2 sc = scale_at(A, (float)tx, (float)ty);
    float b0x, b0y, b1x, b1y;
    basis(t.z, A.su * sc.x, A.sv * sc.y, b0x, b0y, b1x, b1y);
    float ex, ey;
    new_coord(x - (float)tx, y - (float)ty, t.x, t.y, b0x, b0y, b1x, b1y, ex, ey);
    clamp_ex(ex, ey, A.ew, A.eh);
    float ang = screen_to_ex_angle(A, x, y, ex, ey, l);
    csc = scale_at(A, x, y);
    c = make_float4(ex, ey, ang, 0.f);
}

/* calculatePropagateOffset with the pixel's image orientation `io` and
 * scale `csc_px` already fetched; the same operations. */
__device__ __forceinline__ float ex_angle_io(const PtArgs &A, float io, float ex, float ey);
__device__ __forceinline__ void propagate_offset_io(const PtArgs &A, int tx, int ty, float x, float y, float io,
                                                    float2 csc_px, float4 &c, float2 &csc)
{
    int wx = (int)((float)tx - (float)A.W * floorf((float)tx / (float)A.W));
    int wy = (int)((float)ty - (float)A.H * floorf((float)ty / (float)A.H));
    float4 t = A.base.xyth[wy * A.W + wx];
    float2 sc = scale_at(A, (float)tx, (float)ty);
    float b0x, b0y, b1x, b1y;
    basis(t.z, A.su * sc.x, A.sv * sc.y, b0x, b0y, b1x, b1y);
    float ex, ey;
    new_coord(x - (float)tx, y - (float)ty, t.x, t.y, b0x, b0y, b1x, b1y, ex, ey);
    clamp_ex(ex, ey, A.ew, A.eh);
    c = make_float4(ex, ey, ex_angle_io(A, io, ex, ey), 0.f);
    csc = csc_px;
}

/* The keymask of keyframe layer k (texSynth_kernel.cu:1137-1142, 887-892):
 * the exemplar output of that layer at the pixel's own integer position (the
 * texel itself, clamped); alpha > 0 marks a painted keyframe pixel. */
__device__ __forceinline__ bool keymask(const PtArgs &A, int k, int x, int y)
{
    if (k < 0) return false;
    return A.ex[2 * (lay(A, k) + clampi(y, 0, A.eh - 1) * A.ew + clampi(x, 0, A.ew - 1)) + 1].w > 0.f;
}

/* A cumulative key-to-key field texel (k = 0 backward, 1 forward) at
 * exemplar (ex, ey) of layer l: bilinear, clamped in x, y and layer. */
__device__ __forceinline__ float4 cumul_at(const PtArgs &A, int k, float ex, float ey, int l)
{
    int c = l <= 0 ? 0 : (l < A.cl ? l : A.cl - 1);
    return tex4_lin(A.cumul[k] + (size_t)c * A.cw * A.ch, A.cw, A.cw, A.ch, ex, ey);
}

/* The candidates of propagate_kernel_keys / scatter_kernel_keys
 * (texSynth_kernel.cu:1061-1107, 792-834) that follow candidate (ex0, ey0)
 * of layer el through the cumulative fields: to the earlier keys (field 0,
 * layers el-1 down to the first key, each step read in the layer above) and
 * to the later keys (field 1, layers el+1 up to the last key, read in the
 * layer below), each tested in turn with scale (1, 1).  The field texel is
 * Color4(time_step, y, x, layer): propagation stops a walk on a time step
 * below -EPSILON, scatter on an invalid ribbon (layer < 0). */
template <int RT, bool TK>
__device__ void key_walks(const PtArgs &A, const Tile &T, int x, int y, float ex0, float ey0, int el, float ia,
                          bool scatter_rule, float4 &ow, float2 &osc, int &ol, float &best, bool &moved)
{
#pragma unroll 1
    for (int dir = 0; dir < 2; dir++) {
        float ex = ex0, ey = ey0;
        const int step = dir == 0 ? -1 : 1;
#pragma unroll 1
        for (int j = el + step; dir == 0 ? j >= A.key_first : j <= A.key_last; j += step) {
            float4 m = cumul_at(A, dir, ex, ey, j - step);
            if (scatter_rule ? ((int)m.w < 0) : (m.x < -0.000001f)) break;
            ex += m.z;
            ey += m.y;
            float4 t = make_float4(ex, ey, screen_to_ex_angle(A, (float)x, (float)y, ex, ey, j), 0.f);
            float2 ts = make_float2(1.f, 1.f);
            float r = test_offset<RT, TK>(A, T, x, y, t, ts, ia, j);
            if (r < best) {
                ow = t;
                osc = ts;
                ol = j;
                best = r;
                moved = true;
            }
        }
    }
}

/* propagate_kernel (texSynth_kernel.cu:1112-1219) and, in the RT < 0
 * variant, propagate_kernel_keys (:984-1110): style layers, the keymask and
 * the walks through the cumulative fields. */
template <int RT, bool TK>
__device__ __forceinline__ void b_propagate(const PtArgs &A, int bx, int by, int jump)
{
    __shared__ float4 scol[TileSize<RT>::N], scan[TileSize<RT>::N];
    int x, y;
    bool valid = pixel_of(A, bx, by, x, y);
    Tile T = stage_tile<RT>(A, bx, by, scol, scan);
    if (!valid) return;
    int o = y * A.W + x;
    const bool layers = RT < 0 && A.layered;
    if (image_culled(A, x, y)) {
        A.work.xyth[o] = make_float4(0.f, 0.f, 0.f, 0.f);
        A.work.sc[o] = make_float2(1.f, 1.f);
        lum_set(A, o, false);
        if (layers) A.work.la[o] = 0;
        return;
    }
    if (layers && keymask(A, A.key_index, x, y)) {
        A.work.xyth[o] = make_float4((float)x, (float)y, 0.f, 1.f);
        A.work.sc[o] = make_float2(1.f, 1.f);
        lum_set(A, o, false);
        A.work.la[o] = A.key_index;
        return;
    }
    float best = A.res_base[o];
    float ia = A.color[A.cur][o].w;
    float4 ow = A.base.xyth[o];
    float2 osc = A.base.sc[o];
    int ol = layers ? A.base.la[o] : 0;
    bool moved = false;
#pragma unroll 1
    for (int i = 0; i < 4; i++) {
        int tx = x, ty = y;
        if (i == 0) tx = x + jump;
        else if (i == 1) tx = x - jump;
        else if (i == 2) ty = y + jump;
        else ty = y - jump;
        float4 c;
        float2 csc;
        int cl = 0;
        if (layers) propagate_offset(A, tx, ty, (float)x, (float)y, c, csc, &cl);
        else propagate_offset(A, tx, ty, (float)x, (float)y, c, csc);
        float r = test_offset<RT, TK>(A, T, x, y, c, csc, ia, cl);
        if (r < best) {
            ow = c;
            osc = csc;
            ol = cl;
            best = r;
            moved = true;
        }
        if (RT < 0 && A.keys) key_walks<RT, TK>(A, T, x, y, c.x, c.y, cl, ia, false, ow, osc, ol, best, moved);
    }
    A.work.xyth[o] = ow;
    A.work.sc[o] = osc;
    lum_set(A, o, !moved);
    if (layers) A.work.la[o] = ol;
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_BX *PT_BY) void k_propagate(PtArgs A, int jump)
{
    int bx, by;
    xcd_tile(A, bx, by);
    b_propagate<RT, TK>(A, bx, by, jump);
}

__device__ __forceinline__ float u24(uint32_t w) { return (float)(w >> 8) * (1.0f / 16777216.0f); }

/* scatter_kernel (texSynth_kernel.cu:865-952) with averageOffset (:841-862)
 * and, in the RT < 0 variant, scatter_kernel_keys (:723-839): style layers,
 * the keymask, samples in the winner's layer followed through the cumulative
 * fields, no averaged candidate.  As in the reference, a pixel is written
 * only when a candidate beats its cached residual; others keep the working
 * buffer's current content. */
template <int RT, bool TK>
__device__ __forceinline__ void b_scatter(const PtArgs &A, int bx, int by)
{
    __shared__ float4 scol[TileSize<RT>::N], scan[TileSize<RT>::N];
    int x, y;
    bool valid = pixel_of(A, bx, by, x, y);
    Tile T = stage_tile<RT>(A, bx, by, scol, scan);
    if (!valid) return;
    int o = y * A.W + x;
    if (image_culled(A, x, y)) return;
    const bool layers = RT < 0 && A.layered;
    const bool keys = RT < 0 && A.keys;
    float4 off = A.base.xyth[o];
    float2 offsc = A.base.sc[o];
    if (layers && keymask(A, A.key_index, x, y)) { /* keep the current offset */
        A.work.xyth[o] = off;
        A.work.sc[o] = offsc;
        lum_set(A, o, true);
        A.work.la[o] = A.base.la[o];
        return;
    }
    float cached = A.res_base[o];
    float best = (float)((double)cached + 10e-7);
    float4 ow = off;
    float2 osc = offsc;
    int ol = layers ? A.base.la[o] : 0;
    bool moved = false;
    float ia = A.color[A.cur][o].w;
    float dist = (float)(A.ew > A.eh ? A.ew : A.eh);
    const float closest = 10.f, factor = 0.9f;
    float2 sc = scale_at(A, (float)x, (float)y);
    for (int i = 0; (i < A.samples) && (dist * factor >= closest); i++) {
        float r = 0.0f, theta = 0.0f;
        if (A.rng_mode == PT_RNG_COUNTER) {
            uint32_t ctr[4] = {(uint32_t)x, (uint32_t)y, A.call, (uint32_t)i}, key[2] = {A.key0, A.key1}, rnd[4];
            pt_philox4x32_10(ctr, key, rnd);
            r = u24(rnd[0]);
            theta = (float)((double)u24(rnd[1]) * 0.63661977236758134308);
        }
        float dmax = dist, dmin = dist * factor;
        float s, co;
        pt_sincosf(theta, &s, &co);
        float len = r * (dmax - dmin) + dmin;
        float ex = ow.x + co * len;
        float ey = ow.y + s * len;
        clamp_ex(ex, ey, A.ew, A.eh);
        if (keys) {
            /* the sample stays in the winner's layer, scale (1, 1) */
            const int el = ol;
            float4 c = make_float4(ex, ey, screen_to_ex_angle(A, (float)x, (float)y, ex, ey, el), 0.f);
            float2 c1 = make_float2(1.f, 1.f);
            float res = test_offset<RT, TK>(A, T, x, y, c, c1, ia, el);
            if (res < best) {
                ow = c;
                osc = c1;
                best = res;
            }
            key_walks<RT, TK>(A, T, x, y, ex, ey, el, ia, true, ow, osc, ol, best, moved);
        } else {
            float ang = screen_to_ex_angle(A, (float)x, (float)y, ex, ey);
            float4 c = make_float4(ex, ey, ang, 0.f);
            float res = test_offset<RT, TK>(A, T, x, y, c, sc, ia);
            if (res < best) {
                ow = c;
                osc = sc;
                ol = 0;
                best = res;
            }
        }
        dist *= factor;
    }
    if (!keys) {
        /* averageOffset: the neighbours of layer 0 */
        float ax = 0.f, ay = 0.f;
        int sum = 0;
        for (int i = -1; i <= 1; i++)
            for (int j = -1; j <= 1; j++) {
                if (!in_bounds(x + i, y + j, A.lw, A.lh, 0) || (i == 0 && j == 0)) continue;
                if (layers && A.base.la[(y + j) * A.W + (x + i)] != 0) continue;
                float4 n = A.base.xyth[(y + j) * A.W + (x + i)];
                ax += n.x;
                ay += n.y;
                sum++;
            }
        ax /= (float)sum;
        ay /= (float)sum;
        float4 avg = make_float4(ax, ay, screen_to_ex_angle(A, (float)x, (float)y, ax, ay), 1.f);
        float2 avgsc = make_float2(1.f, 1.f);
        float res = test_offset<RT, TK>(A, T, x, y, avg, avgsc, ia);
        if (res < best) {
            ow = avg;
            osc = avgsc;
            ol = 0;
            best = res;
        }
    }
    (void)moved;
    if (best < cached) {
        A.work.xyth[o] = ow;
        A.work.sc[o] = osc;
        lum_set(A, o, false);
        if (layers) A.work.la[o] = ol;
    }
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_BX *PT_BY) void k_scatter(PtArgs A)
{
    int bx, by;
    xcd_tile(A, bx, by);
    b_scatter<RT, TK>(A, bx, by);
}

/* ------------------------------------------------------------------ */
/* exact early rejection of candidates                                 */
/* ------------------------------------------------------------------ */
/*
 * With every weight non-negative and no hysteresis term, each tap adds a
 * non-negative amount to the patch energy, so any partial sum is a lower
 * bound of the reference's in-order total.  A candidate is rejected as soon
 * as a partial sum reaches T = best * sum_w * (1 + 1e-4); the margin exceeds
 * the worst-case relative error of two FP32 recursive sums of (2r+1)^2 <= 225
 * non-negative terms (about 1.6e-5 at r = 7) plus the rounding of T, so a
 * rejected candidate's in-order value r = sum_r / sum_w is provably >= best
 * and the reference's strict `r < best` test would have rejected it too.
 * Survivors are re-evaluated in the reference order (phase 1), giving the
 * bit-identical value the plain kernels compute.
 *
 * Phase 0 evaluates the three heaviest rows (i = 0, -1, +1: ~88% of the
 * Gaussian weight) first; phase 1 runs rows -r..r in order with the same
 * bound check after each row.  Lanes advance through their own candidates
 * row by row, so a wave does not wait on its slowest candidate.
 */
struct Cand {
    float4 t;   /* x, y, theta, hysteresis */
    float2 sc;  /* scaleU, scaleV */
    float b0x, b0y, b1x, b1y;
};

/* Temporal state of the pixel a unit belongs to: the advected offsets and
 * their bases (time_setup) and the sum's start value, minus the hysteresis
 * term when the forward advected offset is valid (calc_residual_impl). */
struct TimeCand {
    float4 pf, pb;
    bool cpv, cnv;
    float tb4[8];
    float acc0;
};

__device__ __forceinline__ float ee_threshold(float best, float sw) { return best * sw * 1.0001f; }

/* Sum of the in-bounds tap weights of pixel (x, y), in the reference order. */
__device__ __forceinline__ float tap_weight_sum(const PtArgs &A, int r, int x, int y)
{
    float sw = 0.f;
    for (int i = -r; i <= r; i++) {
        if (x + i < 0 || x + i >= A.lw) continue;
        for (int j = -r; j <= r; j++) {
            if (y + j < 0 || y + j >= A.lh) continue;
            sw += A.gauss[i + r] * A.gauss[j + r];
        }
    }
    return sw;
}

/* One row (fixed i) of calculateResidual's tap loop, accumulated into
 * (acc, swa) with the reference's per-tap arithmetic; RT < 0 adds the
 * extended terms in calc_residual_impl's order (the tile's target and
 * canvas are staged converted to LAB when that colour space is on). */
template <int RT, bool TIME, bool XL = false, bool TS = false>
__device__ __forceinline__ void row_accum(const PtArgs &A, const Tile &T, __amdgpu_buffer_rsrc_t er, int x, int y,
                                          int i, const Cand &c, float4 pf, float4 pb, bool cpv, bool cnv,
                                          const float *tb4, float &acc, float &swa)
{
    const int r = PT_RAD(RT, A.r);
    const int ew = A.ew, eh = A.eh;
    const float hw = (float)A.hw, hh = (float)A.hh;
    const float gi = A.gauss[i + r];
    const float ibx = (float)i * c.b0x, iby = (float)i * c.b0y;
    const float4 *tcol = T.col + (T.ly + r) * T.tw + (T.lx + r + i);
    const float4 *tcan = T.can + (T.ly + r) * T.tw + (T.lx + r + i);
#pragma unroll
    for (int j = -r; j <= r; j++) {
        if (y + j < 0 || y + j >= A.lh)
            continue;
        float4 base = tcol[j * T.tw];
        float4 canvas = tcan[j * T.tw];
        float tx = c.t.x + (ibx + (float)j * c.b1x);
        float ty = c.t.y + (iby + (float)j * c.b1y);
        tx = wrap_tap(tx, hw);
        ty = wrap_tap(ty, hh);
        float4 tb, to;
        ex_pair_t<XL>(T, er, ew, eh, tx, ty, tb, to);
        if (Rt<RT>::EXT && A.lab) {
            tb = lab4(tb);
            to = lab4(to);
        }
        float res = ssd_alpha(base, tb, A.alpha_w) * A.in_w;
        if (TIME) {
            bool any_time_w = A.tdi > 0.f || A.tdo > 0.f || A.tcw > 0.f;
            bool fwd_dir = A.direction == PT_FORWARD || A.direction == PT_BIDIRECTIONAL;
            bool bwd_dir = A.direction == PT_BACKWARD || A.direction == PT_BIDIRECTIONAL;
            if (TS && T.sdir >= 0) {
                /* one direction in use, its advected set staged in LDS */
                if (any_time_w) {
                    const int si = (T.ly + r + j) * T.tw + (T.lx + r + i);
                    const bool k0 = T.sdir == 0;
                    if ((k0 ? (fwd_dir && cpv) : (bwd_dir && cnv)) && T.av[si]) {
                        const float *b4 = tb4 + (k0 ? 0 : 4);
                        res = res + time_terms_v<Rt<RT>::EXT, true>(A, er, T.sdir, k0 ? pf : pb, b4[0], b4[1],
                                                                    b4[2], b4[3], base, tb, canvas, to, T.ab[si],
                                                                    T.ao[si], x, y, i, j);
                    }
                }
            } else if (any_time_w) {
                int oo = (y + j) * A.W + (x + i);
                if (fwd_dir && cpv && A.adv[0].xyth[oo].w > 0)
                    res = res + time_terms_b<Rt<RT>::EXT>(A, er, 0, pf, tb4[0], tb4[1], tb4[2], tb4[3], base, tb,
                                                          canvas, to, oo, x, y, i, j);
                if (bwd_dir && cnv && A.adv[1].xyth[oo].w > 0)
                    res = res + time_terms_b<Rt<RT>::EXT>(A, er, 1, pb, tb4[4], tb4[5], tb4[6], tb4[7], base, tb,
                                                          canvas, to, oo, x, y, i, j);
            }
        }
        if (i == 0 && j == 0) {
            if (A.dt_w > 0.f) {
                float de = tex1_lin(A.ex_dt, A.ew, A.ew, A.eh, tx, ty);
                de /= c.sc.y;
                float di = A.dt[A.cur][y * A.W + x];
                float diff = fabsf(de - di);
                diff = fminf(fmaxf(diff, 0.0f), 10.0f);
                res = res + diff * A.dt_w;
            }
            if (Rt<RT>::EXT && A.obj_w > 0.f) {
                int iid = A.img_id[point_index(A.idw, A.idh, (float)x, (float)y)];
                int sid = A.ex_id[point_index(A.ew, A.eh, tx, ty)];
                if (iid != sid) res = res + A.obj_w;
            }
        } else {
            res = fmaf(ssd_alpha(canvas, to, A.alpha_w), A.canvas_w, res);
            if (Rt<RT>::EXT) res = coherence_add(A, res, x, y, i, j, c.t.x, c.t.y, tx, ty);
        }
        float wgt = gi * A.gauss[j + r];
        acc = fmaf(res, wgt, acc);
        swa += wgt;
    }
}

/* row_accum split in two for a fixed radius: the tap residuals of row i
 * (no temporal terms) ... */
template <int RT, bool XL = false>
__device__ __forceinline__ void row_taps(const PtArgs &A, const Tile &T, __amdgpu_buffer_rsrc_t er, int x, int y, int i,
                                         const Cand &c, float (&rv)[2 * RT + 1])
{
    constexpr int r = RT;
    const int ew = A.ew, eh = A.eh;
    const float hw = (float)A.hw, hh = (float)A.hh;
    const float ibx = (float)i * c.b0x, iby = (float)i * c.b0y;
    const float4 *tcol = T.col + (T.ly + r) * T.tw + (T.lx + r + i);
    const float4 *tcan = T.can + (T.ly + r) * T.tw + (T.lx + r + i);
#pragma unroll
    for (int j = -r; j <= r; j++) {
        rv[j + r] = 0.f;
        if (y + j < 0 || y + j >= A.lh)
            continue;
        float4 base = tcol[j * T.tw];
        float4 canvas = tcan[j * T.tw];
        float tx = c.t.x + (ibx + (float)j * c.b1x);
        float ty = c.t.y + (iby + (float)j * c.b1y);
        tx = wrap_tap(tx, hw);
        ty = wrap_tap(ty, hh);
        float4 tb, to;
        ex_pair_t<XL>(T, er, ew, eh, tx, ty, tb, to);
        float res = ssd_alpha(base, tb, A.alpha_w) * A.in_w;
        if (i == 0 && j == 0) {
            if (A.dt_w > 0.f) {
                float de = tex1_lin(A.ex_dt, A.ew, A.ew, A.eh, tx, ty);
                de /= c.sc.y;
                float di = A.dt[A.cur][y * A.W + x];
                float diff = fabsf(de - di);
                diff = fminf(fmaxf(diff, 0.0f), 10.0f);
                res = res + diff * A.dt_w;
            }
        } else {
            res = fmaf(ssd_alpha(canvas, to, A.alpha_w), A.canvas_w, res);
        }
        rv[j + r] = res;
    }
}

/* ... and their accumulation in the reference order, with the same skips
 * and weights as row_accum. */
template <int RT>
__device__ __forceinline__ void row_sum(const PtArgs &A, int y, int i, const float (&rv)[2 * RT + 1], float &acc,
                                        float &swa)
{
    constexpr int r = RT;
    const float gi = A.gauss[i + r];
#pragma unroll
    for (int j = -r; j <= r; j++) {
        if (y + j < 0 || y + j >= A.lh)
            continue;
        float wgt = gi * A.gauss[j + r];
        acc = fmaf(rv[j + r], wgt, acc);
        swa += wgt;
    }
}

/* Temporal bases of the advected offsets at (x, y) (texSynth_kernel.cu:552-557). */
template <bool TIME>
__device__ __forceinline__ void time_setup(const PtArgs &A, int x, int y, float4 &pf, float4 &pb, bool &cpv,
                                           bool &cnv, float *tb4)
{
    for (int k = 0; k < 8; k++) tb4[k] = 0.f;
    cpv = cnv = false;
    pf = pb = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!TIME) return;
    int o = y * A.W + x;
    pf = A.adv[0].xyth[o];
    pb = A.adv[1].xyth[o];
    cpv = pf.w > 0;
    cnv = pb.w > 0;
    if (cpv) {
        float2 ps = A.adv[0].sc[o];
        float ang = screen_to_ex_angle(A, (float)x, (float)y, pf.x, pf.y);
        basis(ang, A.su * ps.x, A.sv * ps.y, tb4[0], tb4[1], tb4[2], tb4[3]);
    }
    if (cnv) {
        float2 ps = A.adv[1].sc[o];
        float ang = screen_to_ex_angle(A, (float)x, (float)y, pb.x, pb.y);
        basis(ang, A.su * ps.x, A.sv * ps.y, tb4[4], tb4[5], tb4[6], tb4[7]);
    }
}

/* Candidate acceptance test of test_offset (texSynth_kernel.cu:695-721)
 * before the patch distance: false means the candidate scores 1e31. */
__device__ __forceinline__ bool candidate_testable(const PtArgs &A, float4 t, float image_alpha)
{
    if (exemplar_culled(A, t.x, t.y)) return false;
    if (!A.transp) {
        float ea = tex_ex_out_alpha(A, t.x, t.y);
        if (ea < 0.99f && image_alpha > 0.99f) return false;
    }
    return true;
}

/*
 * Runs one lane's candidate stream through the two-phase evaluation.  NEXT
 * produces candidate k (returns false when there are no more) and may read
 * the current winner; ACCEPT records a winner.  `best` is updated in place.
 */
template <int RT, bool TIME, class Next, class Accept>
__device__ __forceinline__ void ee_search(const PtArgs &A, const Tile &T, int x, int y, float &best, Next next,
                                          Accept accept)
{
    const int r = PT_RAD(RT, A.r);
    __amdgpu_buffer_rsrc_t er = ex_rsrc(A);
    float4 pf, pb;
    bool cpv, cnv;
    float tb4[8];
    time_setup<TIME>(A, x, y, pf, pb, cpv, cnv, tb4);
    const float sw = tap_weight_sum(A, r, x, y);
    float T0 = ee_threshold(best, sw);
    Cand c;
    int ph = 0, step = 0;
    float acc = 0.f, swa = 0.f;
    bool need = true;
    while (true) {
        if (need) {
            if (!next(c)) break;
            if (!(best > 0.f)) continue; /* nothing can score below a non-positive best */
            basis(c.t.z, A.su * c.sc.x, A.sv * c.sc.y, c.b0x, c.b0y, c.b1x, c.b1y);
            ph = 0;
            step = 0;
            acc = 0.f;
            swa = 0.f;
            need = false;
        }
        int i = ph == 0 ? (step == 0 ? 0 : (step == 1 ? -1 : 1)) : step - r;
        if (x + i >= 0 && x + i < A.lw)
            row_accum<RT, TIME>(A, T, er, x, y, i, c, pf, pb, cpv, cnv, tb4, acc, swa);
        if (acc >= T0) {
            need = true;
        } else if (ph == 0 && step == 2) {
            ph = 1;
            step = 0;
            acc = 0.f;
            swa = 0.f;
        } else if (ph == 1 && step == 2 * r) {
            float rr = (swa < 0.000001f) ? 1e31f : acc / swa;
            if (rr < best) {
                best = rr;
                T0 = ee_threshold(best, sw);
                accept(c);
            }
            need = true;
        } else {
            step++;
        }
    }
}

/* propagate_kernel (texSynth_kernel.cu:1112-1219) with exact early rejection */
template <int RT, bool TK>
__device__ __forceinline__ void b_propagate_ee(const PtArgs &A, int bx, int by, int jump)
{
    __shared__ float4 scol[TileSize<RT>::N], scan[TileSize<RT>::N];
    int x, y;
    bool valid = pixel_of(A, bx, by, x, y);
    Tile T = stage_tile<RT>(A, bx, by, scol, scan);
    if (!valid) return;
    int o = y * A.W + x;
    if (image_culled(A, x, y)) {
        A.work.xyth[o] = make_float4(0.f, 0.f, 0.f, 0.f);
        A.work.sc[o] = make_float2(1.f, 1.f);
        lum_set(A, o, false);
        return;
    }
    float best = A.res_base[o];
    float ia = A.color[A.cur][o].w;
    float4 ow = A.base.xyth[o];
    float2 osc = A.base.sc[o];
    bool moved = false;
    int kk = 0;
    auto next_seq = [&](Cand &c) -> bool {
        while (kk < 4) {
            int k = kk++;
            int tx = x, ty = y;
            if (k == 0) tx = x + jump;
            else if (k == 1) tx = x - jump;
            else if (k == 2) ty = y + jump;
            else ty = y - jump;
            propagate_offset(A, tx, ty, (float)x, (float)y, c.t, c.sc);
            if (candidate_testable(A, c.t, ia)) return true;
        }
        return false;
    };
    auto accept = [&](const Cand &c) {
        ow = c.t;
        osc = c.sc;
        moved = true;
    };
    if (TK) {
        float4 pf = A.adv[0].xyth[o], pb = A.adv[1].xyth[o];
        if (pf.w > 0 || pb.w > 0)
            ee_search<RT, true>(A, T, x, y, best, next_seq, accept);
        else
            ee_search<RT, false>(A, T, x, y, best, next_seq, accept);
    } else {
        ee_search<RT, false>(A, T, x, y, best, next_seq, accept);
    }
    A.work.xyth[o] = ow;
    A.work.sc[o] = osc;
    lum_set(A, o, !moved);
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_BX *PT_BY) void k_propagate_ee(PtArgs A, int jump)
{
    b_propagate_ee<RT, TK>(A, blockIdx.x, blockIdx.y, jump);
}

/* scatter_kernel (texSynth_kernel.cu:865-952) with exact early rejection.
 * Sample k is generated only after samples < k are decided, from the winner
 * so far, exactly as in the reference's sequential loop. */
template <int RT, bool TK>
__device__ __forceinline__ void b_scatter_ee(const PtArgs &A, int bx, int by)
{
    __shared__ float4 scol[TileSize<RT>::N], scan[TileSize<RT>::N];
    int x, y;
    bool valid = pixel_of(A, bx, by, x, y);
    Tile T = stage_tile<RT>(A, bx, by, scol, scan);
    if (!valid) return;
    int o = y * A.W + x;
    if (image_culled(A, x, y)) return;
    float cached = A.res_base[o];
    float best = (float)((double)cached + 10e-7);
    float4 ow = A.base.xyth[o];
    float2 osc = A.base.sc[o];
    float ia = A.color[A.cur][o].w;
    const float factor = 0.9f;
    float dist = (float)(A.ew > A.eh ? A.ew : A.eh);
    float2 sc = scale_at(A, (float)x, (float)y);
    int i = 0;          /* next random sample */
    bool avg_done = false;
    auto next = [&](Cand &c) -> bool {
        while (i < A.samples && dist * factor >= 10.f) {
            float r = 0.0f, theta = 0.0f;
            if (A.rng_mode == PT_RNG_COUNTER) {
                uint32_t ctr[4] = {(uint32_t)x, (uint32_t)y, A.call, (uint32_t)i}, key[2] = {A.key0, A.key1}, rnd[4];
                pt_philox4x32_10(ctr, key, rnd);
                r = u24(rnd[0]);
                theta = (float)((double)u24(rnd[1]) * 0.63661977236758134308);
            }
            float dmax = dist, dmin = dist * factor;
            float s, co;
            pt_sincosf(theta, &s, &co);
            float len = r * (dmax - dmin) + dmin;
            float ex = ow.x + co * len;
            float ey = ow.y + s * len;
            clamp_ex(ex, ey, A.ew, A.eh);
            float ang = screen_to_ex_angle(A, (float)x, (float)y, ex, ey);
            c.t = make_float4(ex, ey, ang, 0.f);
            c.sc = sc;
            dist *= factor;
            i++;
            if (candidate_testable(A, c.t, ia)) return true;
        }
        if (!avg_done) {
            avg_done = true;
            float ax = 0.f, ay = 0.f;
            int sum = 0;
            for (int a = -1; a <= 1; a++)
                for (int b = -1; b <= 1; b++) {
                    if (!in_bounds(x + a, y + b, A.lw, A.lh, 0) || (a == 0 && b == 0)) continue;
                    float4 n = A.base.xyth[(y + b) * A.W + (x + a)];
                    ax += n.x;
                    ay += n.y;
                    sum++;
                }
            ax /= (float)sum;
            ay /= (float)sum;
            c.t = make_float4(ax, ay, screen_to_ex_angle(A, (float)x, (float)y, ax, ay), 1.f);
            c.sc = make_float2(1.f, 1.f);
            if (candidate_testable(A, c.t, ia)) return true;
        }
        return false;
    };
    auto accept = [&](const Cand &c) {
        ow = c.t;
        osc = c.sc;
    };
    if (TK) {
        float4 pf = A.adv[0].xyth[o], pb = A.adv[1].xyth[o];
        if (pf.w > 0 || pb.w > 0)
            ee_search<RT, true>(A, T, x, y, best, next, accept);
        else
            ee_search<RT, false>(A, T, x, y, best, next, accept);
    } else {
        ee_search<RT, false>(A, T, x, y, best, next, accept);
    }
    if (best < cached) {
        A.work.xyth[o] = ow;
        A.work.sc[o] = osc;
        lum_set(A, o, false);
    }
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_BX *PT_BY) void k_scatter_ee(PtArgs A)
{
    b_scatter_ee<RT, TK>(A, blockIdx.x, blockIdx.y);
}

/* screenToExemplarAngle with the image orientation at the pixel already
 * fetched (io = imgOrient(x, y)); the same operations. */
__device__ __forceinline__ float ex_angle_io(const PtArgs &A, float io, float ex, float ey)
{
    if (A.orient_en) {
        float theta = A.ex_orient[point_index(A.ew, A.eh, ex, ey)] - io;
        if ((double)theta > 1.57079632679489661923)
            return (float)((double)theta - 3.14159265358979323846);
        if ((double)theta < -1.57079632679489661923)
            return (float)(3.14159265358979323846 + (double)theta);
        return theta;
    }
    return 0.0f;
}

__device__ __forceinline__ float image_orient_at(const PtArgs &A, int x, int y)
{
    return A.orient[point_index(A.ow, A.oh, (float)x, (float)y)];
}

/* Scatter sample k of pixel (x, y) placed from winner `ow`
 * (texSynth_kernel.cu:909-925); dist is the loop's dist at iteration k. */
__device__ __forceinline__ void scatter_pos(const PtArgs &A, int x, int y, int k, float dist, float4 ow, float &ex,
                                            float &ey)
{
    float rr = 0.0f, theta = 0.0f;
    if (A.rng_mode == PT_RNG_COUNTER) {
        uint32_t ctr[4] = {(uint32_t)x, (uint32_t)y, A.call, (uint32_t)k}, key[2] = {A.key0, A.key1}, rnd[4];
        pt_philox4x32_10(ctr, key, rnd);
        rr = (float)(rnd[0] >> 8) * (1.0f / 16777216.0f);
        theta = (float)((double)((float)(rnd[1] >> 8) * (1.0f / 16777216.0f)) * 0.63661977236758134308);
    }
    float dmax = dist, dmin = dist * 0.9f;
    float sn, co;
    pt_sincosf(theta, &sn, &co);
    float len = rr * (dmax - dmin) + dmin;
    ex = ow.x + co * len;
    ey = ow.y + sn * len;
    clamp_ex(ex, ey, A.ew, A.eh);
}

__device__ __forceinline__ float4 scatter_sample(const PtArgs &A, int x, int y, int k, float dist, float4 ow, float io)
{
    float ex, ey;
    scatter_pos(A, x, y, k, dist, ow, ex, ey);
    return make_float4(ex, ey, ex_angle_io(A, io, ex, ey), 0.f);
}

/* averageOffset (texSynth_kernel.cu:841-862) with its angle (:938-940). */
__device__ __forceinline__ float4 average_candidate(const PtArgs &A, int x, int y, float io)
{
    float ax = 0.f, ay = 0.f;
    int sum = 0;
    for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++) {
            if (!in_bounds(x + a, y + b, A.lw, A.lh, 0) || (a == 0 && b == 0)) continue;
            float4 nb = A.base.xyth[(y + b) * A.W + (x + a)];
            ax += nb.x;
            ay += nb.y;
            sum++;
        }
    ax /= (float)sum;
    ay /= (float)sum;
    return make_float4(ax, ay, ex_angle_io(A, io, ax, ay), 1.f);
}

/* ------------------------------------------------------------------ */
/* unit-parallel candidate search                                      */
/* ------------------------------------------------------------------ */
/*
 * A workgroup owns a TX x TY tile of pixels.  Every candidate of every pixel
 * is a unit; lanes pull units from an LDS queue, evaluate them (with the
 * exact early rejection of ee_search, or fully), and store each unit's value
 * (+inf when it provably cannot beat the pixel's best at the time).  A
 * per-pixel pass then applies the reference's sequential rule, candidate by
 * candidate, `if (r < best) { winner = candidate; best = r; }`.
 *
 * Propagate candidates do not depend on earlier acceptances, so one round
 * suffices.  Scatter sample k is placed relative to the winner so far, so
 * samples are evaluated speculatively from the current winner; when sample k
 * is accepted, samples after k are evaluated again from the new winner in a
 * further round.  The averaged-neighbour candidate does not depend on the
 * winner and is evaluated in the first round.  The outcome is the reference's
 * sequential result, bit for bit.
 */
#define PT_UP_THREADS 256
#define PT_UP_MAX_SAMPLES 32
#define PT_INF __builtin_inff()

template <int RT, int TX, int TY> struct UpTile {
    static constexpr int R = PT_RAD(RT, PT_MAX_R);
    static constexpr int N = (TX + 2 * R) * (TY + 2 * R);
};

/* stage_tile's staging for the unit kernels (LAB-converted in the
 * extended variants when that colour space is on) */
template <int RT, int TX, int TY>
__device__ __forceinline__ void stage_tile_up(const PtArgs &A, int x0, int y0, float4 *scol, float4 *scan)
{
    const int r = PT_RAD(RT, A.r);
    const int tw = TX + 2 * r, th = TY + 2 * r;
    const float4 *color = A.color[A.cur];
    const bool lab = Rt<RT>::EXT && A.lab;
    for (int k = threadIdx.x; k < tw * th; k += PT_UP_THREADS) {
        int tx = clampi(x0 - r + k % tw, 0, A.lw - 1);
        int ty = clampi(y0 - r + k / tw, 0, A.lh - 1);
        float4 c = color[ty * A.W + tx], v = A.canvas[ty * A.W + tx];
        if (lab) {
            c = lab4(c);
            v = lab4(v);
        }
        scol[k] = c;
        scan[k] = v;
    }
}

/* The advected set of the launch's only temporal direction (forward or
 * backward; -1 when both or neither are in use) over the tile and its r
 * halo, coordinates clamped into the level as stage_tile_up's (taps outside
 * the level are skipped, so clamped entries are never read): validity, and
 * for valid entries the previous-frame colour and exemplar output. */
template <int RT, int TX, int TY>
__device__ __forceinline__ int stage_adv_up(const PtArgs &A, int x0, int y0, float4 *sab, float4 *sao,
                                            unsigned char *sav)
{
    const bool fwd = A.direction == PT_FORWARD || A.direction == PT_BIDIRECTIONAL;
    const bool bwd = A.direction == PT_BACKWARD || A.direction == PT_BIDIRECTIONAL;
    if (fwd == bwd) return -1;
    const int k = fwd ? 0 : 1;
    const PtAdv &d = A.adv[k];
    const int r = PT_RAD(RT, A.r);
    const int tw = TX + 2 * r, th = TY + 2 * r;
    const bool lab = Rt<RT>::EXT && A.lab; /* staged converted (time_terms_v's PREV_LAB) */
    for (int e = threadIdx.x; e < tw * th; e += PT_UP_THREADS) {
        int tx = clampi(x0 - r + e % tw, 0, A.lw - 1);
        int ty = clampi(y0 - r + e / tw, 0, A.lh - 1);
        int o = ty * A.W + tx;
        bool v = d.xyth[o].w > 0;
        sav[e] = v ? 1 : 0;
        if (v) {
            float4 b = d.base[o], u = d.out[o];
            if (lab) {
                b = lab4(b);
                u = lab4(u);
            }
            sab[e] = b;
            sao[e] = u;
        }
    }
    return k;
}

/* Evaluates one unit (candidate c with its basis) for pixel (x, y) at tile
 * position (lx, ly): the exact patch distance, or +inf when a partial sum
 * proves it cannot beat `best` (EE). */
template <int RT, bool EE, int TX, bool TM, bool XL = false>
__device__ __forceinline__ float up_eval(const PtArgs &A, const Tile &T, __amdgpu_buffer_rsrc_t er, int x, int y,
                                         Cand c, const TimeCand &tc, float best, float sw)
{
    const int r = PT_RAD(RT, A.r);
    const float a0 = TM ? tc.acc0 : 0.f;
    float acc = a0, swa = 0.f;
    if (EE) {
        const float T0 = ee_threshold(best, sw);
        /* phase 0: the three heaviest rows */
#pragma unroll 1
        for (int step = 0; step < 3; step++) {
            int i = step == 0 ? 0 : (step == 1 ? -1 : 1);
            if (i > r) break;
            if (x + i >= 0 && x + i < A.lw)
                row_accum<RT, TM, XL, TM>(A, T, er, x, y, i, c, tc.pf, tc.pb, tc.cpv, tc.cnv, tc.tb4, acc, swa);
            if (acc >= T0) return PT_INF;
        }
        acc = a0;
        swa = 0.f;
#pragma unroll 1
        for (int i = -r; i <= r; i++) {
            if (x + i >= 0 && x + i < A.lw)
                row_accum<RT, TM, XL, TM>(A, T, er, x, y, i, c, tc.pf, tc.pb, tc.cpv, tc.cnv, tc.tb4, acc, swa);
            if (acc >= T0) return PT_INF;
        }
    } else {
#pragma unroll 1
        for (int i = -r; i <= r; i++) {
            if (x + i >= 0 && x + i < A.lw)
                row_accum<RT, TM, XL, TM>(A, T, er, x, y, i, c, tc.pf, tc.pb, tc.cpv, tc.cnv, tc.tb4, acc, swa);
        }
    }
    return (swa < 0.000001f) ? 1e31f : acc / swa;
}

/*
 * Row-granular state machine over a queue of units (EE variant): a lane
 * pulls the next unit as soon as its current one is decided, so a wave's
 * time follows the mean cost of its units rather than the slowest one.
 * SETUP(u, c, x, y, p) prepares unit u, basis included (returns false when
 * the unit needs no evaluation; it has then stored its value itself);
 * DONE(u, value) stores a decided unit's value (+inf when rejected).
 */
template <int RT, int TX, bool KEEP, bool TM, bool XL = false, class Setup, class Done>
__device__ __forceinline__ void up_queue_ee(const PtArgs &A, Tile &T, __amdgpu_buffer_rsrc_t er, int first_u,
                                            int U, int *s_next, const float *s_best, const float *s_sw, Setup setup,
                                            Done done)
{
    static_assert(!(KEEP && TM), "the replayed rows carry no temporal terms");
    const int r = PT_RAD(RT, A.r);
    float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float tb4[8];
    Cand c;
    TimeCand tc;
    int u = first_u, x = 0, y = 0, p = 0, ph = 0, step = 0;
    float acc = 0.f, swa = 0.f, T0 = 0.f;
    bool need = true;
    /* KEEP: the tap residuals of rows 0, -1, +1 from the first phase, so the
     * in-order phase replays them instead of gathering them again */
    constexpr int NK = KEEP ? 2 * RT + 1 : 1;
    float k0[NK], km[NK], kp[NK];
    while (true) {
        if (need) {
            bool found = false;
            while (u < U) {
                if (setup(u, c, tc, x, y, p)) {
                    found = true;
                    break;
                }
                u = atomicAdd(s_next, 1);
            }
            if (!found) break;
            float best = s_best[p];
            T0 = ee_threshold(best, s_sw[p]);
            T.lx = p % TX;
            T.ly = p / TX;
            ph = 0;
            step = 0;
            acc = TM ? tc.acc0 : 0.f;
            swa = 0.f;
            need = false;
        }
        int i = ph == 0 ? (step == 0 ? 0 : (step == 1 ? -1 : 1)) : step - r;
        if constexpr (KEEP) {
            if (x + i >= 0 && x + i < A.lw) {
                if (ph == 0) {
                    float rv[NK];
                    row_taps<RT, XL>(A, T, er, x, y, i, c, rv);
                    row_sum<RT>(A, y, i, rv, acc, swa);
#pragma unroll
                    for (int j = 0; j < NK; j++) {
                        k0[j] = step == 0 ? rv[j] : k0[j];
                        km[j] = step == 1 ? rv[j] : km[j];
                        kp[j] = step == 2 ? rv[j] : kp[j];
                    }
                } else if (i >= -1 && i <= 1) {
                    float rv[NK];
#pragma unroll
                    for (int j = 0; j < NK; j++) rv[j] = i == 0 ? k0[j] : (i < 0 ? km[j] : kp[j]);
                    row_sum<RT>(A, y, i, rv, acc, swa);
                } else {
                    row_accum<RT, false, XL>(A, T, er, x, y, i, c, z4, z4, false, false, tb4, acc, swa);
                }
            }
        } else if constexpr (TM) {
            if (x + i >= 0 && x + i < A.lw)
                row_accum<RT, true, XL, true>(A, T, er, x, y, i, c, tc.pf, tc.pb, tc.cpv, tc.cnv, tc.tb4, acc, swa);
        } else {
            if (x + i >= 0 && x + i < A.lw)
                row_accum<RT, false, XL>(A, T, er, x, y, i, c, z4, z4, false, false, tb4, acc, swa);
        }
        float val = 0.f;
        bool decided = false;
        if (acc >= T0) {
            val = PT_INF;
            decided = true;
        } else if (ph == 0 && step == 2) {
            ph = 1;
            step = 0;
            acc = TM ? tc.acc0 : 0.f;
            swa = 0.f;
        } else if (ph == 1 && step == 2 * r) {
            val = (swa < 0.000001f) ? 1e31f : acc / swa;
            decided = true;
        } else {
            step++;
        }
        if (decided) {
            done(u, val);
            need = true;
            u = atomicAdd(s_next, 1);
        }
    }
}

/* Whether the temporal terms of this launch still gather the advected
 * offsets' exemplar samples per candidate (a direction in use has no time
 * table).  Those gathers are shared by the candidates of one pixel, so the
 * unit kernels then order units pixel-major; otherwise candidate-major, where
 * neighbouring lanes take the same candidate of neighbouring pixels (coherent
 * exemplar lines, and neighbouring time-table entries).  The order changes
 * no result. */
template <bool TM> __device__ __forceinline__ bool up_pixel_major(const PtArgs &A)
{
    if (!TM) return false;
    const bool fwd = A.direction == PT_FORWARD || A.direction == PT_BIDIRECTIONAL;
    const bool bwd = A.direction == PT_BACKWARD || A.direction == PT_BIDIRECTIONAL;
    return (fwd && !A.tt[0]) || (bwd && !A.tt[1]);
}

/* propagate_kernel (texSynth_kernel.cu:1112-1219), unit-parallel. */
/* Whether a unit kernel must evaluate candidates of a pixel: it is inside
 * the level and not culled (its tap weight sum is positive), and a
 * candidate could score below its best.  Without a positive hysteresis
 * weight every residual is >= 0, so a best <= 0 (or NaN) cannot be beaten. */
__device__ __forceinline__ bool up_open(const PtArgs &A, float sw, float best)
{
    return sw > 0.f && (best > 0.f || A.hyst_w > 0.f);
}

/* Per-pixel temporal state kept in LDS by the unit kernels. */
template <int P> struct UpTime {
    float2 pf[P], pb[P];
    float tb4[P][8];
    float acc0[P];
    int flags[P]; /* bit 0: forward advected offset valid, bit 1: backward */
};

template <int P>
__device__ __forceinline__ void up_time_init(const PtArgs &A, UpTime<P> &S, int p, int x, int y, bool live)
{
    float4 pf, pb;
    bool cpv = false, cnv = false;
    float tb4[8];
    if (live) time_setup<true>(A, x, y, pf, pb, cpv, cnv, tb4);
    else {
        pf = pb = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int k = 0; k < 8; k++) tb4[k] = 0.f;
    }
    S.pf[p] = make_float2(pf.x, pf.y);
    S.pb[p] = make_float2(pb.x, pb.y);
    for (int k = 0; k < 8; k++) S.tb4[p][k] = tb4[k];
    float a0 = 0.f;
    if (cpv) a0 -= hysteresis_term(A, x, y);
    S.acc0[p] = a0;
    S.flags[p] = (cpv ? 1 : 0) | (cnv ? 2 : 0);
}

template <int P> __device__ __forceinline__ void up_time_load(const UpTime<P> &S, int p, TimeCand &tc)
{
    tc.pf = make_float4(S.pf[p].x, S.pf[p].y, 0.f, 0.f);
    tc.pb = make_float4(S.pb[p].x, S.pb[p].y, 0.f, 0.f);
    for (int k = 0; k < 8; k++) tc.tb4[k] = S.tb4[p][k];
    tc.acc0 = S.acc0[p];
    tc.cpv = (S.flags[p] & 1) != 0;
    tc.cnv = (S.flags[p] & 2) != 0;
}

/*
 * Two phases: every lane first sets up candidate units in lockstep (the
 * neighbour's offset carried over, its angle, the acceptance test, the
 * patch basis), storing them in LDS and listing the testable ones; then the
 * lanes pull listed units through the row state machine, whose per-unit
 * setup is a few LDS reads, so a wave does not pay the full setup whenever
 * one of its lanes starts a new unit.
 */
/* Exemplar window side of the XL propagate: 24 texels (18 KB) hold the
 * patches of a coherent 8x8 tile's candidates while keeping four workgroups
 * per CU (measured on the 1024^2 still: 20 / 24 / 28 / 32 texels give
 * 120.9 / 119.7 / 122.6 / 123.8 ms per frame). */
#define PT_XW_PROP 24
/* the temporal variants' window: 20 texels, so that their LDS (the staged
 * advected set on top) leaves room for four workgroups per CU */
#define PT_XW_PROP_TM 20
template <bool TM> struct PropXw {
    static constexpr int W = TM ? PT_XW_PROP_TM : PT_XW_PROP;
};

template <int RT, int TX, int TY, bool TM, bool XL> struct PropUpLds {
    static constexpr int P = TX * TY, U = 4 * P, N = UpTile<RT, TX, TY>::N;
    UpTime<TM ? P : 1> s_tm;
    /* XL: a window of the interleaved exemplar anchored on the current offset
     * of the tile's middle pixel; units whose every tap footprint lies in it
     * read their taps from LDS (same texels, weights and order) */
    float4 s_exw[XL ? PropXw<TM>::W * PropXw<TM>::W * 2 : 1];
    unsigned char s_exl[XL ? U : 1];
    /* TM: the advected set of the launch's direction over the tile */
    float4 s_ab[TM ? N : 1], s_ao[TM ? N : 1];
    unsigned char s_av[TM ? N : 1];
    float4 scol[N], scan[N];
    float s_best[P], s_sw[P], s_ia[P], s_io[P];
    float2 s_sc[P];
    float s_r[U];
    float s_cx[U], s_cy[U], s_b0x[U], s_b0y[U], s_b1x[U], s_b1y[U], s_scy[U];
    unsigned short s_list[U];
    int s_n, s_nb, s_next;
};

template <int RT, bool EE, int TX, int TY, bool TM, bool XL = false, bool AR = false>
__device__ __forceinline__ void b_propagate_up(const PtArgs &A, int bx, int by, int jump, void *arena = nullptr)
{
    constexpr int P = TX * TY, U = 4 * P;
    static_assert(!XL || Rt<RT>::R > 0, "exemplar windows need a fixed radius");
    using Lds = PropUpLds<RT, TX, TY, TM, XL>;
    constexpr int N = Lds::N;
    PT_LDS_VAR(UpTime<TM ? P : 1>, s_tm)
    PT_LDS_ARRAY(float4, s_exw, (XL ? PropXw<TM>::W * PropXw<TM>::W * 2 : 1))
    PT_LDS_ARRAY(unsigned char, s_exl, (XL ? U : 1))
    PT_LDS_ARRAY(float4, s_ab, (TM ? N : 1))
    PT_LDS_ARRAY(float4, s_ao, (TM ? N : 1))
    PT_LDS_ARRAY(unsigned char, s_av, (TM ? N : 1))
    PT_LDS_ARRAY(float4, scol, N)
    PT_LDS_ARRAY(float4, scan, N)
    PT_LDS_ARRAY(float, s_best, P)
    PT_LDS_ARRAY(float, s_sw, P)
    PT_LDS_ARRAY(float, s_ia, P)
    PT_LDS_ARRAY(float, s_io, P)
    PT_LDS_ARRAY(float2, s_sc, P)
    PT_LDS_ARRAY(float, s_r, U)
    PT_LDS_ARRAY(float, s_cx, U)
    PT_LDS_ARRAY(float, s_cy, U)
    PT_LDS_ARRAY(float, s_b0x, U)
    PT_LDS_ARRAY(float, s_b0y, U)
    PT_LDS_ARRAY(float, s_b1x, U)
    PT_LDS_ARRAY(float, s_b1y, U)
    PT_LDS_ARRAY(float, s_scy, U)
    PT_LDS_ARRAY(unsigned short, s_list, U)
    PT_LDS_VAR(int, s_n)
    PT_LDS_VAR(int, s_nb)
    PT_LDS_VAR(int, s_next)
    const int x0 = bx * TX, y0 = by * TY;
    const int r = PT_RAD(RT, A.r);
    stage_tile_up<RT, TX, TY>(A, x0, y0, scol, scan);
    int sdir = -1;
    if constexpr (TM) sdir = stage_adv_up<RT, TX, TY>(A, x0, y0, s_ab, s_ao, s_av);
    int wx = 0, wy = 0;
    if constexpr (XL) {
        int ax = min(x0 + TX / 2, A.lw - 1), ay = min(y0 + TY / 2, A.lh - 1);
        float4 an = A.base.xyth[ay * A.W + ax];
        stage_window(A, an.x, an.y, PropXw<TM>::W, s_exw, wx, wy, threadIdx.x, PT_UP_THREADS);
    }
    for (int p = threadIdx.x; p < P; p += PT_UP_THREADS) {
        int x = x0 + p % TX, y = y0 + p / TX;
        bool live = x < A.lw && y < A.lh && !image_culled(A, x, y);
        int o = y * A.W + x;
        s_best[p] = live ? A.res_base[o] : -1.f;
        s_sw[p] = live ? tap_weight_sum(A, r, x, y) : 0.f;
        s_ia[p] = live ? A.color[A.cur][o].w : 0.f;
        s_io[p] = live ? image_orient_at(A, x, y) : 0.f;
        s_sc[p] = live ? scale_at(A, (float)x, (float)y) : make_float2(1.f, 1.f);
        if constexpr (TM) up_time_init<P>(A, s_tm, p, x, y, live);
    }
    if (threadIdx.x == 0) {
        s_n = 0;
        s_nb = 0;
        s_next = PT_UP_THREADS;
    }
    __syncthreads();
    /* phase 1: candidate set-up in lockstep, in the order of up_pixel_major
     * (the list, and the first units each lane evaluates, follow it) */
    const bool pmaj = up_pixel_major<TM>(A) || (TM && (A.time_unit_order == 2 || (XL && A.time_unit_order == 1)));
    for (int v = threadIdx.x; v < U; v += PT_UP_THREADS) {
        int p = pmaj ? v >> 2 : v % P, k = pmaj ? v & 3 : v / P, u = 4 * p + k;
        int x = x0 + p % TX, y = y0 + p / TX;
        s_r[u] = PT_INF;
        if (!up_open(A, s_sw[p], s_best[p])) continue;
        int tx = x, ty = y;
        if (k == 0) tx = x + jump;
        else if (k == 1) tx = x - jump;
        else if (k == 2) ty = y + jump;
        else ty = y - jump;
        float4 c;
        float2 csc;
        propagate_offset_io(A, tx, ty, (float)x, (float)y, s_io[p], s_sc[p], c, csc);
        /* while the residual cache is exact (A.res_exact), a candidate equal
         * to the pixel's current offset (position, angle and scale) has the
         * cached residual bit for bit, the best it must beat strictly: it is
         * not evaluated */
        if (A.res_exact) {
            const int o = y * A.W + x;
            const float4 own = A.base.xyth[o];
            const float2 osc = A.base.sc[o];
            if (c.x == own.x && c.y == own.y && c.z == own.z && csc.x == osc.x && csc.y == osc.y) continue;
        }
        if (!candidate_testable(A, c, s_ia[p])) continue;
        float b0x, b0y, b1x, b1y;
        basis(c.z, A.su * csc.x, A.sv * csc.y, b0x, b0y, b1x, b1y);
        if constexpr (XL) s_exl[u] = in_window(A, r, c.x, c.y, b0x, b0y, b1x, b1y, wx, wy, PropXw<TM>::W) ? 1 : 0;
        s_cx[u] = c.x;
        s_cy[u] = c.y;
        s_b0x[u] = b0x;
        s_b0y[u] = b0y;
        s_b1x[u] = b1x;
        s_b1y[u] = b1y;
        s_scy[u] = csc.y;
        if constexpr (XL) {
            /* units inside the window from the front of the list, the others
             * from its end, so that a wave mostly takes units of one kind
             * and issues one kind of gather per tap */
            if (s_exl[u]) s_list[atomicAdd(&s_n, 1)] = (unsigned short)u;
            else s_list[U - 1 - atomicAdd(&s_nb, 1)] = (unsigned short)u;
        } else {
            s_list[atomicAdd(&s_n, 1)] = (unsigned short)u;
        }
    }
    __syncthreads();
    /* phase 2: evaluation of the listed units (list position of unit index
     * idx: the window units, then the others from the end backwards) */
    const int na = s_n;
    const int n = XL ? na + s_nb : na;
    auto at = [&](int idx) { return (!XL || idx < na) ? idx : U - 1 - (idx - na); };
    __amdgpu_buffer_rsrc_t er = ex_rsrc(A);
    Tile T;
    T.col = scol;
    T.can = scan;
    T.tw = TX + 2 * r;
    T.exw = s_exw;
    T.wx = wx;
    T.wy = wy;
    T.ww = PropXw<TM>::W;
    T.exl = false;
    T.ab = s_ab;
    T.ao = s_ao;
    T.av = s_av;
    T.sdir = sdir;
    auto setup = [&](int idx, Cand &c, TimeCand &tc, int &x, int &y, int &p) -> bool {
        int u = s_list[at(idx)];
        p = u >> 2;
        x = x0 + p % TX;
        y = y0 + p / TX;
        if constexpr (XL) T.exl = s_exl[u] != 0;
        c.t = make_float4(s_cx[u], s_cy[u], 0.f, 0.f);
        c.sc = make_float2(1.f, s_scy[u]);
        c.b0x = s_b0x[u];
        c.b0y = s_b0y[u];
        c.b1x = s_b1x[u];
        c.b1y = s_b1y[u];
        if constexpr (TM) up_time_load<P>(s_tm, p, tc);
        return true;
    };
    auto done = [&](int idx, float v) { s_r[s_list[at(idx)]] = v; };
    if (EE) {
        up_queue_ee<RT, TX, (RT > 0) && !TM, TM, XL>(A, T, er, threadIdx.x, n, &s_next, s_best, s_sw, setup, done);
    } else {
        for (int idx = threadIdx.x; idx < n;) {
            Cand c;
            TimeCand tc;
            int x, y, p;
            setup(idx, c, tc, x, y, p);
            T.lx = p % TX;
            T.ly = p / TX;
            done(idx, up_eval<RT, false, TX, TM, XL>(A, T, er, x, y, c, tc, s_best[p], s_sw[p]));
            idx = atomicAdd(&s_next, 1);
        }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < P; p += PT_UP_THREADS) {
        int x = x0 + p % TX, y = y0 + p / TX;
        if (x >= A.lw || y >= A.lh) continue;
        int o = y * A.W + x;
        if (image_culled(A, x, y)) {
            A.work.xyth[o] = make_float4(0.f, 0.f, 0.f, 0.f);
            A.work.sc[o] = make_float2(1.f, 1.f);
            lum_set(A, o, false);
            continue;
        }
        float best = A.res_base[o];
        int win = -1;
        for (int k = 0; k < 4; k++) {
            float v = s_r[4 * p + k];
            if (v < best) {
                best = v;
                win = k;
            }
        }
        lum_set(A, o, win < 0);
        if (win < 0) {
            A.work.xyth[o] = A.base.xyth[o];
            A.work.sc[o] = A.base.sc[o];
        } else {
            int tx = x, ty = y;
            if (win == 0) tx = x + jump;
            else if (win == 1) tx = x - jump;
            else if (win == 2) ty = y + jump;
            else ty = y - jump;
            float4 c;
            float2 csc;
            propagate_offset_io(A, tx, ty, (float)x, (float)y, s_io[p], s_sc[p], c, csc);
            A.work.xyth[o] = c;
            A.work.sc[o] = csc;
        }
    }
}

template <int RT, bool EE, int TX, int TY, bool TM, bool XL = false>
__global__ __launch_bounds__(PT_UP_THREADS) __attribute__((amdgpu_waves_per_eu(TM ? 4 : 1))) void
k_propagate_up(PtArgs A, int jump)
{
    int bx, by;
    xcd_tile(A, bx, by);
    b_propagate_up<RT, EE, TX, TY, TM, XL>(A, bx, by, jump);
}

/* scatter_kernel (texSynth_kernel.cu:865-952), unit-parallel in rounds.
 * n = number of random samples (the reference's loop bound, identical for
 * every pixel); unit k == n is the averaged-neighbour candidate. */
template <int RT, int TX, int TY, bool TM> struct ScatUpLds {
    static constexpr int P = TX * TY, S = PT_UP_MAX_SAMPLES + 1, N = UpTile<RT, TX, TY>::N;
    UpTime<TM ? P : 1> s_tm;
    float4 scol[N], scan[N];
    float4 s_ow[P];
    float2 s_osc[P];
    float s_best[P], s_sw[P], s_ia[P], s_cached[P], s_iori[P];
    float2 s_sc[P];
    int s_start[P];
    float s_dist[S];
    float s_r[P * S];
    int s_next, s_more;
    float4 s_ab[TM ? N : 1], s_ao[TM ? N : 1];
    unsigned char s_av[TM ? N : 1];
};

template <int RT, bool EE, int TX, int TY, bool TM, bool AR = false>
__device__ __forceinline__ void b_scatter_up(const PtArgs &A, int bx, int by, int n, void *arena = nullptr)
{
    constexpr int P = TX * TY;
    constexpr int S = PT_UP_MAX_SAMPLES + 1;
    using Lds = ScatUpLds<RT, TX, TY, TM>;
    constexpr int N = Lds::N;
    PT_LDS_VAR(UpTime<TM ? P : 1>, s_tm)
    PT_LDS_ARRAY(float4, scol, N)
    PT_LDS_ARRAY(float4, scan, N)
    PT_LDS_ARRAY(float4, s_ow, P)
    PT_LDS_ARRAY(float2, s_osc, P)
    PT_LDS_ARRAY(float, s_best, P)
    PT_LDS_ARRAY(float, s_sw, P)
    PT_LDS_ARRAY(float, s_ia, P)
    PT_LDS_ARRAY(float, s_cached, P)
    PT_LDS_ARRAY(float, s_iori, P)
    PT_LDS_ARRAY(float2, s_sc, P)
    PT_LDS_ARRAY(int, s_start, P)
    PT_LDS_ARRAY(float, s_dist, S)
    PT_LDS_ARRAY(float, s_r, P * S)
    PT_LDS_VAR(int, s_next)
    PT_LDS_VAR(int, s_more)
    PT_LDS_ARRAY(float4, s_ab, (TM ? N : 1))
    PT_LDS_ARRAY(float4, s_ao, (TM ? N : 1))
    PT_LDS_ARRAY(unsigned char, s_av, (TM ? N : 1))
    const int x0 = bx * TX, y0 = by * TY;
    const int r = PT_RAD(RT, A.r);
    const int SU = n + 1; /* units per pixel */
    const bool pmaj = up_pixel_major<TM>(A) || (TM && A.time_unit_order == 2);
    stage_tile_up<RT, TX, TY>(A, x0, y0, scol, scan);
    int sdir = -1;
    if constexpr (TM) sdir = stage_adv_up<RT, TX, TY>(A, x0, y0, s_ab, s_ao, s_av);
    if (threadIdx.x == 0) {
        float d = (float)(A.ew > A.eh ? A.ew : A.eh);
        for (int k = 0; k < n; k++) {
            s_dist[k] = d;
            d *= 0.9f;
        }
    }
    for (int p = threadIdx.x; p < P; p += PT_UP_THREADS) {
        int x = x0 + p % TX, y = y0 + p / TX;
        bool live = x < A.lw && y < A.lh && !image_culled(A, x, y);
        int o = y * A.W + x;
        float cached = live ? A.res_base[o] : 0.f;
        s_cached[p] = cached;
        s_best[p] = live ? (float)((double)cached + 10e-7) : -1.f;
        s_sw[p] = live ? tap_weight_sum(A, r, x, y) : 0.f;
        s_ia[p] = live ? A.color[A.cur][o].w : 0.f;
        s_ow[p] = live ? A.base.xyth[o] : make_float4(0.f, 0.f, 0.f, 0.f);
        s_osc[p] = live ? A.base.sc[o] : make_float2(1.f, 1.f);
        s_iori[p] = live ? image_orient_at(A, x, y) : 0.f;
        s_sc[p] = live ? scale_at(A, (float)x, (float)y) : make_float2(1.f, 1.f);
        s_start[p] = live ? 0 : SU;
        if constexpr (TM) up_time_init<P>(A, s_tm, p, x, y, live);
    }
    __amdgpu_buffer_rsrc_t er = ex_rsrc(A);
    Tile T;
    T.col = scol;
    T.can = scan;
    T.tw = TX + 2 * r;
    T.exl = false;
    T.ab = s_ab;
    T.ao = s_ao;
    T.av = s_av;
    T.sdir = sdir;
    const float2 one = make_float2(1.f, 1.f);
    bool first = true;
    while (true) {
        if (threadIdx.x == 0) s_next = PT_UP_THREADS;
        __syncthreads();
        /* unit u: pixel p, sample k, sample-major or pixel-major as
         * up_pixel_major says */
        auto setup = [&](int u, Cand &c, TimeCand &tc, int &x, int &y, int &p) -> bool {
            p = pmaj ? u / SU : u % P;
            int k = pmaj ? u % SU : u / P;
            int st = s_start[p];
            bool todo = (k < n) ? (k >= st && st < n) : (first && st <= n);
            if (!todo) return false;
            x = x0 + p % TX;
            y = y0 + p / TX;
            if (k < n) {
                c.t = scatter_sample(A, x, y, k, s_dist[k], s_ow[p], s_iori[p]);
                c.sc = s_sc[p];
                if (A.rng_mode == PT_RNG_REFERENCE && k > st) {
                    /* the truncated reference draw places every sample on the
                     * +x ray from the winner: samples past the exemplar edge
                     * all clamp to its last column.  A sample equal to the
                     * previous one of this round (same winner, same scale)
                     * has that one's residual bit for bit, so it cannot be the
                     * first to beat the best: it is not evaluated. */
                    float px, py;
                    scatter_pos(A, x, y, k - 1, s_dist[k - 1], s_ow[p], px, py);
                    if (px == c.t.x && py == c.t.y) {
                        s_r[p * S + k] = PT_INF;
                        return false;
                    }
                }
            } else {
                c.t = average_candidate(A, x, y, s_iori[p]);
                c.sc = one;
            }
            if (!up_open(A, s_sw[p], s_best[p]) || !candidate_testable(A, c.t, s_ia[p])) {
                s_r[p * S + k] = PT_INF;
                return false;
            }
            basis(c.t.z, A.su * c.sc.x, A.sv * c.sc.y, c.b0x, c.b0y, c.b1x, c.b1y);
            if constexpr (TM) up_time_load<P>(s_tm, p, tc);
            return true;
        };
        auto done = [&](int u, float v) { s_r[pmaj ? (u / SU) * S + (u % SU) : (u % P) * S + (u / P)] = v; };
        if (EE) {
            up_queue_ee<RT, TX, false, TM>(A, T, er, threadIdx.x, P * SU, &s_next, s_best, s_sw, setup, done);
        } else {
            for (int u = threadIdx.x; u < P * SU;) {
                Cand c;
                TimeCand tc;
                int x, y, p;
                if (setup(u, c, tc, x, y, p)) {
                    T.lx = p % TX;
                    T.ly = p / TX;
                    done(u, up_eval<RT, false, TX, TM>(A, T, er, x, y, c, tc, s_best[p], s_sw[p]));
                }
                u = atomicAdd(&s_next, 1);
            }
        }
        if (threadIdx.x == 0) s_more = 0;
        __syncthreads();
        /* resolution of the samples of this round, in order */
        for (int p = threadIdx.x; p < P; p += PT_UP_THREADS) {
            int st = s_start[p];
            if (st >= n) continue;
            int x = x0 + p % TX, y = y0 + p / TX;
            float best = s_best[p];
            int k = st;
            for (; k < n; k++) {
                float v = s_r[p * S + k];
                if (v < best) break;
            }
            if (k < n) {
                /* sample k wins: rebuild it from the winner it was placed from */
                s_ow[p] = scatter_sample(A, x, y, k, s_dist[k], s_ow[p], s_iori[p]);
                s_osc[p] = s_sc[p];
                s_best[p] = s_r[p * S + k];
                s_start[p] = k + 1;
                if (k + 1 < n) s_more = 1;
            } else {
                s_start[p] = n;
            }
        }
        first = false;
        __syncthreads();
        if (!s_more) break;
    }
    /* averaged-neighbour candidate, then the write-back rule */
    for (int p = threadIdx.x; p < P; p += PT_UP_THREADS) {
        int x = x0 + p % TX, y = y0 + p / TX;
        if (x >= A.lw || y >= A.lh || image_culled(A, x, y)) continue;
        int o = y * A.W + x;
        float best = s_best[p];
        float4 ow = s_ow[p];
        float2 osc = s_osc[p];
        float v = s_r[p * S + n];
        if (v < best) {
            best = v;
            ow = average_candidate(A, x, y, s_iori[p]);
            osc = one;
        }
        if (best < s_cached[p]) {
            A.work.xyth[o] = ow;
            A.work.sc[o] = osc;
            lum_set(A, o, false);
        }
    }
}

template <int RT, bool EE, int TX, int TY, bool TM>
__global__ __launch_bounds__(PT_UP_THREADS) void k_scatter_up(PtArgs A, int n)
{
    int bx, by;
    xcd_tile(A, bx, by);
    b_scatter_up<RT, EE, TX, TY, TM>(A, bx, by, n);
}

/* upsampleBilinearKernel (texSynth_kernel.cu:1453-1510) with
 * offsetUpsampleNearest (:1403-1417). */
__device__ __forceinline__ void b_upsample(const PtArgs &A, int bx, int by)
{
    int x, y;
    if (!pixel_of_rows(A, bx, by, x, y)) return;
    if (!in_bounds(x, y, A.lw, A.lh, 1)) return;
    float cxf = (float)((double)x - 0.5) / 2.0f;
    float cyf = (float)((double)y - 0.5) / 2.0f;
    int lcx = (int)floorf(cxf), lcy = (int)floorf(cyf);
    float4 o00 = A.base.xyth[lcy * A.W + lcx];
    float4 o10 = A.base.xyth[lcy * A.W + lcx + 1];
    float4 o01 = A.base.xyth[(lcy + 1) * A.W + lcx];
    float4 o11 = A.base.xyth[(lcy + 1) * A.W + lcx + 1];
    float a = cxf - (float)lcx, b = cyf - (float)lcy;
    float w00 = (1 - a) * (1 - b), w10 = a * (1 - b), w01 = (1 - a) * b, w11 = a * b;
    float ix = ((w00 * o00.x + w10 * o10.x) + w01 * o01.x) + w11 * o11.x;
    float iy = ((w00 * o00.y + w10 * o10.y) + w01 * o01.y) + w11 * o11.y;
    float ux = ix * 2.0f + 0.5f, uy = iy * 2.0f + 0.5f;
    const float M = 4.0f * 4.0f;
#define D2(q) (((ix - (q).x) * (ix - (q).x)) + ((iy - (q).y) * (iy - (q).y)))
    if (D2(o00) > M || D2(o10) > M || D2(o01) > M || D2(o11) > M) {
        int rcx = (int)roundf(cxf), rcy = (int)roundf(cyf);
        float4 cx = A.base.xyth[rcy * A.W + rcx];
        float b0x, b0y, b1x, b1y;
        basis(cx.z, 1, 1, b0x, b0y, b1x, b1y);
        float nx, ny;
        new_coord(cxf - (float)rcx, cyf - (float)rcy, cx.x, cx.y, b0x, b0y, b1x, b1y, nx, ny);
        clamp_ex(nx, ny, A.ew, A.eh);
        ux = (float)(((double)nx * 2.0) + 0.5);
        uy = (float)(((double)ny * 2.0) + 0.5);
    }
#undef D2
    float rw = A.res_base[lcy * A.W + lcx];
    /* the style layer of the top-left coarse offset */
    int l00 = A.layered ? A.base.la[lcy * A.W + lcx] : 0;
    float ang = screen_to_ex_angle(A, (float)x, (float)y, ux, uy, l00);
    float2 sc = scale_at(A, (float)x, (float)y);
    int o = y * A.W + x;
    A.work.xyth[o] = make_float4(ux, uy, ang, 0.f);
    A.work.sc[o] = sc;
    lum_set(A, o, false);
    if (A.layered) A.work.la[o] = l00;
    A.res_work[o] = rw;
}

__global__ __launch_bounds__(PT_BX *PT_BY) void k_upsample(PtArgs A) { b_upsample(A, blockIdx.x, blockIdx.y); }

/* calculateTranslatedPosition (texSynth_kernel.cu:1541-1575) */
__device__ __forceinline__ bool translated_position(const PtArgs &A, int fwd, int x, int y, float &tx, float &ty)
{
    int prev = (A.cur + 1) % 2;
    float2 forward, back;
    if (fwd) {
        back = tex2_lin(A.velb[A.cur], A.W, A.W, A.H, (float)x, (float)y);
        tx = (float)x + back.x;
        ty = (float)y + back.y;
        forward = tex2_lin(A.velf[prev], A.W, A.W, A.H, tx, ty);
    } else {
        forward = tex2_lin(A.velf[A.cur], A.W, A.W, A.H, (float)x, (float)y);
        tx = (float)x + forward.x;
        ty = (float)y + forward.y;
        back = tex2_lin(A.velb[prev], A.W, A.W, A.H, tx, ty);
    }
    const float valid_dist = 1.414 * 2;
    forward.x = -forward.x;
    forward.y = -forward.y;
    float dist = sqrtf((forward.x - back.x) * (forward.x - back.x) + (forward.y - back.y) * (forward.y - back.y));
    return in_bounds((int)ceilf(tx), (int)ceilf(ty), A.lw, A.lh, 0) &&
           in_bounds((int)floorf(tx), (int)floorf(ty), A.lw, A.lh, 0) && (dist < valid_dist);
}

/* A frame-to-key ribbon texel (k = 0 backward, 1 forward) at the integer
 * pixel (x, y): the texel itself, clamped.  Color4(time_step, y, x, layer). */
__device__ __forceinline__ float4 ftk_at(const PtArgs &A, int k, int x, int y)
{
    return A.ftk[k][clampi(y, 0, A.fh[k] - 1) * A.fw[k] + clampi(x, 0, A.fw[k] - 1)];
}

/* advectKernel (texSynth_kernel.cu:1577-1693) */
__global__ __launch_bounds__(PT_BX *PT_BY) void k_advect(PtArgs A, int fwd, int time_step, int flags)
{
    int x, y;
    if (!pixel_of_rows(A, blockIdx.x, blockIdx.y, x, y)) return;
    int o = y * A.W + x;
    const PtAdv &d = A.adv[fwd ? 0 : 1];
    if (image_culled(A, x, y)) {
        d.base[o] = make_float4(0, 0, 0, 0);
        d.out[o] = make_float4(0, 0, 0, 0);
        d.xyth[o] = make_float4(0, 0, 0, 0);
        d.sc[o] = make_float2(1.f, 1.f);
        if (A.lum_live) d.lum[o] = 0.f;
        if (A.layered) d.la[o] = 0;
        return;
    }
    int prev = (A.cur + 1) % 2;
    float4 rnd = A.random[point_index(A.rw, A.rh, (float)x, (float)y)];
    float4 frac = make_float4(rnd.x - floorf(rnd.x), rnd.y - floorf(rnd.y), rnd.z - floorf(rnd.z),
                              rnd.w - floorf(rnd.w));
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    float lum = 0.f;
    int ol = 0;
    float tx, ty;
    /* isImagePixelCulledPreviousFrame(tx, ty, level, 8.0) takes int
     * coordinates (texSynth_kernel.cu:188-198): the position is truncated and
     * the texel read exactly; translated_position has put it in bounds. */
    if (!(flags & PT_ADVECT_DO_NOT_CHECK_PREVIOUS) && translated_position(A, fwd, x, y, tx, ty) &&
        !(A.dt[prev][(int)ty * A.W + (int)tx] > ((float)A.cull - 8.0f))) {
        int rtx = (int)roundf(tx), rty = (int)roundf(ty);
        float4 nr = A.base.xyth[rty * A.W + rtx];
        if (A.layered) ol = A.base.la[rty * A.W + rtx];
        float b0x, b0y, b1x, b1y;
        basis(nr.z, 1, 1, b0x, b0y, b1x, b1y);
        float ix, iy;
        new_coord(tx - (float)rtx, ty - (float)rty, nr.x, nr.y, b0x, b0y, b1x, b1y, ix, iy);
        clamp_ex(ix, iy, A.ew, A.eh);
        float4 prev_color = tex4_lin(A.color[prev], A.W, A.W, A.H, tx, ty);
        out = make_float4(ix, iy, nr.z, 1.f);
        if (A.lum_live) {
            /* luminanceShift: the nearest source's shift plus the luminance
             * change of the image between the frames (:1628-1631) */
            float4 col = A.color[A.cur][o];
            lum = A.lum_enable ? A.base.lum[rty * A.W + rtx] + pt_luminance_delta(col.x, col.y, col.z, prev_color.x,
                                                                                    prev_color.y, prev_color.z)
                               : 0.f;
        }
        d.base[o] = prev_color;
        d.out[o] = tex_ex_out(A, out.x, out.y, ol);
        d.ts[o] = time_step;
    }
    /* ADVECT_CHECK_FRAME_TO_KEY (:1643-1668): a long step, or a failed
     * advection, may take the pixel's ribbon to a keyframe instead */
    if ((flags & PT_ADVECT_CHECK_FRAME_TO_KEY) && (time_step > 1 || out.w == 0.f)) {
        float4 fk = ftk_at(A, fwd ? 0 : 1, x, y);
        int fts = (int)fk.x, fla = (int)fk.w;
        if (fla >= 0 && (abs(fts) < time_step || out.w == 0.f)) {
            out.x = fk.z + (float)x;
            out.y = fk.y + (float)y;
            ol = fla;
            out.w = 1.f;
            lum = 0.f;
        }
        float4 eb, eo;
        ex_pair(ex_rsrc(A), A.ew, A.eh, out.x, out.y, eb, eo, lay(A, ol));
        d.base[o] = eb;
        d.out[o] = eo;
        d.ts[o] = abs(fts);
    }
    if (out.w == 0.f) {
        out = make_float4(rnd.x, rnd.y, rnd.z, 0.f);
        ol = (int)rnd.w;
        lum = 0.f;
        d.base[o] = make_float4(0, 0, 0, 0);
        d.out[o] = make_float4(0, 0, 0, 0);
        d.ts[o] = -1;
    }
    float jit = A.jitter;
    out.x = out.x * (1 - jit) + frac.x * (jit);
    out.y = out.y * (1 - jit) + frac.y * (jit);
    out.z = screen_to_ex_angle(A, (float)x, (float)y, out.x, out.y, ol);
    d.xyth[o] = out;
    d.sc[o] = make_float2(1.f, 1.f);
    if (A.lum_live) d.lum[o] = lum;
    if (A.layered) d.la[o] = ol;
}

/* The candidate-independent exemplar samples of timeDerivativeTerms
 * (texSynth_kernel.cu:448-452) for advected set k: for every pixel whose
 * advected offset is valid, the exemplar base and output at each tap of that
 * offset's window, with the basis and coordinates of time_setup /
 * calc_residual_impl (same operations, so the same values).  Tap-major
 * planes: for a fixed tap, neighbouring pixels are neighbouring entries. */
__global__ __launch_bounds__(PT_RX *PT_RY) void k_time_table(PtArgs A, int k, float4 *tt)
{
    int x, y;
    if (!pixel_of_rows(A, blockIdx.x, blockIdx.y, x, y)) return;
    const int o = y * A.W + x;
    const float4 prev = A.adv[k].xyth[o];
    if (!(prev.w > 0)) return; /* never read: the centre offset is not valid */
    const float2 ps = A.adv[k].sc[o];
    float b0x, b0y, b1x, b1y;
    float ang = screen_to_ex_angle(A, (float)x, (float)y, prev.x, prev.y);
    basis(ang, A.su * ps.x, A.sv * ps.y, b0x, b0y, b1x, b1y);
    __amdgpu_buffer_rsrc_t er = ex_rsrc(A);
    const int r = A.r;
    for (int i = -r; i <= r; i++) {
        for (int j = -r; j <= r; j++) {
            float ptx, pty;
            new_coord((float)i, (float)j, prev.x, prev.y, b0x, b0y, b1x, b1y, ptx, pty);
            float4 peb, peo;
            ex_pair(er, A.ew, A.eh, ptx, pty, peb, peo);
            if (A.lab) { /* read by the extended variants only */
                peb = lab4(peb);
                peo = lab4(peo);
            }
            float4 *e = tt + time_table_index(A, x, y, i, j);
            e[0] = peb;
            e[1] = peo;
        }
    }
}

/* mergeRandomlyKernel (texSynth_kernel.cu:1891-1936): the advected offset
 * also wins where this direction's frame-to-key ribbon is valid and the
 * other one is not (a context without ribbons binds zero texels, valid
 * layer-0 ribbons, which leave that clause false). */
__global__ __launch_bounds__(PT_BX *PT_BY) void k_merge_randomly(PtArgs A, int fwd, float weight, int force)
{
    int x, y;
    if (!pixel_of_rows(A, blockIdx.x, blockIdx.y, x, y)) return;
    int o = y * A.W + x;
    const PtAdv &d = A.adv[fwd ? 0 : 1];
    /* both sources' offsets are read before the choice (one round of
     * loads: the launch is latency-bound, every pixel's loads are in bounds) */
    float4 a = d.xyth[o], b = A.base.xyth[o];
    float2 asc = d.sc[o], bsc = A.base.sc[o];
    float4 rnd = A.random[point_index(A.rw, A.rh, (float)x, (float)y)];
    float r = fwd ? rnd.x - floorf(rnd.x) : rnd.y - floorf(rnd.y);
    bool rib = (int)ftk_at(A, fwd ? 0 : 1, x, y).w >= 0 && !((int)ftk_at(A, fwd ? 1 : 0, x, y).w >= 0);
    if (rib || ((a.w > 0 || force) && r < weight)) {
        A.work.xyth[o] = a;
        A.work.sc[o] = asc;
        if (A.lum_live) A.work.lum[o] = d.lum[o];
        if (A.layered) A.work.la[o] = d.la[o];
    } else {
        A.work.xyth[o] = b;
        A.work.sc[o] = bsc;
        lum_set(A, o, true);
        if (A.layered) A.work.la[o] = A.base.la[o];
    }
}

/* mergeUsingResidualKernel (texSynth_kernel.cu:1856-1889) */
__global__ __launch_bounds__(PT_BX *PT_BY) void k_merge_residual(PtArgs A, int fwd)
{
    int x, y;
    if (!pixel_of_rows(A, blockIdx.x, blockIdx.y, x, y)) return;
    int o = y * A.W + x;
    const PtAdv &d = A.adv[fwd ? 0 : 1];
    if (d.res[o] < A.res_base[o]) {
        float4 a = d.xyth[o];
        a.w = 1.f;
        A.work.xyth[o] = a;
        A.work.sc[o] = d.sc[o];
        if (A.lum_live) A.work.lum[o] = d.lum[o];
        if (A.layered) A.work.la[o] = d.la[o];
    } else {
        float4 u = A.base.xyth[o];
        u.w = 0.f;
        A.work.xyth[o] = u;
        A.work.sc[o] = A.base.sc[o];
        lum_set(A, o, true);
        if (A.layered) A.work.la[o] = A.base.la[o];
    }
}

/* mergeAdvectedKernel (texSynth_kernel.cu:1938-1976): both advected offsets
 * valid -> one of them by the fractional random x, one valid -> it, none ->
 * the fractional random texel as an offset (MERGE_FALLBACK_RANDOM) or the
 * current base offset. */
__global__ __launch_bounds__(PT_BX *PT_BY) void k_merge_advected(PtArgs A, int fallback)
{
    int x, y;
    if (!pixel_of_rows(A, blockIdx.x, blockIdx.y, x, y)) return;
    int o = y * A.W + x;
    float4 f = A.adv[0].xyth[o], b = A.adv[1].xyth[o];
    float4 rnd = A.random[point_index(A.rw, A.rh, (float)x, (float)y)];
    /* randomFloat4's out parameter: the fractional parts (:207-213) */
    float4 fr = make_float4(rnd.x - floorf(rnd.x), rnd.y - floorf(rnd.y), rnd.z - floorf(rnd.z),
                            rnd.w - floorf(rnd.w));
    int k = -1; /* advected set taken, -1: none */
    if (f.w > 0 && b.w > 0) k = (fr.x < 0.5f) ? 0 : 1;
    else if (f.w > 0) k = 0;
    else if (b.w > 0) k = 1;
    if (k >= 0) {
        const PtAdv &d = A.adv[k];
        A.work.xyth[o] = d.xyth[o];
        A.work.sc[o] = d.sc[o];
        if (A.lum_live) A.work.lum[o] = d.lum[o];
        A.work.la[o] = d.la[o];
    } else if (fallback == PT_MERGE_FALLBACK_RANDOM) {
        A.work.xyth[o] = make_float4(fr.x, fr.y, fr.z, 0.f);
        A.work.sc[o] = make_float2(1.f, 1.f);
        lum_set(A, o, false);
        A.work.la[o] = (int)fr.w;
    } else {
        A.work.xyth[o] = A.base.xyth[o];
        A.work.sc[o] = A.base.sc[o];
        lum_set(A, o, true);
        A.work.la[o] = A.base.la[o];
    }
}

/* copyFrameToKeyRibbonKernel (texSynth_kernel.cu:2061-2097): each pixel
 * starts at the keyframe texel its ribbon reaches (hysteresis 1, scale
 * (1, 1)), or at its random texel where the ribbon is invalid. */
__global__ __launch_bounds__(PT_BX *PT_BY) void k_offsets_from_ribbon(PtArgs A, int look_forwards)
{
    int x, y;
    if (!pixel_of_rows(A, blockIdx.x, blockIdx.y, x, y)) return;
    int o = y * A.W + x;
    float4 rb = ftk_at(A, look_forwards ? 1 : 0, x, y);
    int rl = (int)rb.w;
    if (rl >= 0) {
        float ex = rb.z + (float)x, ey = rb.y + (float)y;
        A.work.xyth[o] = make_float4(ex, ey, screen_to_ex_angle(A, (float)x, (float)y, ex, ey, rl), 1.f);
        A.work.la[o] = rl;
    } else {
        float4 r = A.random[point_index(A.rw, A.rh, (float)x, (float)y)];
        A.work.xyth[o] = make_float4(r.x, r.y, r.z, 0.f);
        A.work.la[o] = (int)r.w;
    }
    A.work.sc[o] = make_float2(1.f, 1.f);
    lum_set(A, o, false);
}

__device__ __forceinline__ bool ribbon_valid(const pt_ribbonp &r) { return r.layer >= 0; }

/* accumulateAdvectionKernel (texSynth_kernel.cu:1728-1826): the ribbon
 * field `src` advected one step of time_step frames along the image
 * velocities into `dst`.  Where the four source ribbons around the
 * translated position are all valid, of one layer and within 5 px of each
 * other the position is interpolated bilinearly, else the nearest one is
 * taken; the step's displacement and time are then added.  Without
 * `overwrite` a valid ribbon already in dst is replaced only by one of no
 * longer time. */
__global__ __launch_bounds__(PT_BX *PT_BY) void k_accumulate_ribbon(PtArgs A, const pt_ribbonp *src, pt_ribbonp *dst,
                                                                   int fwd, int time_step, int overwrite)
{
    int x, y;
    if (!pixel_of_rows(A, blockIdx.x, blockIdx.y, x, y)) return;
    int o = y * A.W + x;
    pt_ribbonp out = {0.f, 0.f, 0, -1};
    float tx, ty;
    if (A.color[A.cur][o].w == 0.f || !translated_position(A, fwd, x, y, tx, ty)) {
        if (overwrite) dst[o] = out;
        return;
    }
    int ftx = (int)floorf(tx), fty = (int)floorf(ty), ctx = (int)ceilf(tx), cty = (int)ceilf(ty);
    pt_ribbonp s00 = src[fty * A.W + ftx], s10 = src[fty * A.W + ctx];
    pt_ribbonp s01 = src[cty * A.W + ftx], s11 = src[cty * A.W + ctx];
    bool v00 = ribbon_valid(s00), v10 = ribbon_valid(s10), v01 = ribbon_valid(s01), v11 = ribbon_valid(s11);
    if (!(v00 || v10 || v01 || v11)) {
        if (overwrite) dst[o] = out;
        return;
    }
    bool all_valid = v00 && v10 && v01 && v11;
    /* the reference's chained comparison of the four layers, evaluated as
     * C++ does: the bool of one comparison compared with the next layer */
    bool same = ((int)((int)(s00.layer == s10.layer) == s01.layer)) == s11.layer;
    const float th = 5.f;
    bool diverges = fabsf(s00.x - s10.x) > th || fabsf(s00.x - s01.x) > th || fabsf(s00.x - s11.x) > th ||
                    fabsf(s00.y - s10.y) > th || fabsf(s00.y - s01.y) > th || fabsf(s00.y - s11.y) > th;
    if (!all_valid || !same || diverges) {
        int rtx = (int)roundf(tx), rty = (int)roundf(ty);
        if (ftx == rtx && fty == rty) out = s00;
        else if (ftx + 1 == rtx && fty == rty) out = s10;
        else if (ftx == rtx && fty + 1 == rty) out = s01;
        else out = s11;
    } else {
        float wx = tx - (float)ftx, wy = ty - (float)fty;
        float rx0 = s00.x * (1.f - wx) + s10.x * wx;
        float ry0 = s00.y * (1.f - wx) + s10.y * wx;
        float rx1 = s01.x * (1.f - wx) + s11.x * wx;
        float ry1 = s01.y * (1.f - wx) + s11.y * wx;
        out.x = rx0 * (1.f - wy) + rx1 * wy;
        out.y = ry0 * (1.f - wy) + ry1 * wy;
        out.time_step = s00.time_step;
        out.layer = s00.layer;
    }
    out.x += (tx - (float)x);
    out.y += (ty - (float)y);
    out.time_step += time_step;
    pt_ribbonp cur = dst[o];
    if (overwrite || !ribbon_valid(cur) || fabsf((float)cur.time_step) >= fabsf((float)out.time_step)) dst[o] = out;
}

/* cumulativeAdvectionFieldToOutput (texSynth_kernel.cu:290-305): the
 * canvas texel each valid ribbon reaches (clamped into the level-0-sized
 * canvas) as the forward advected output colour, zero elsewhere. */
__global__ __launch_bounds__(PT_BX *PT_BY) void k_ribbon_output(PtArgs A, const pt_ribbonp *rib)
{
    int x, y;
    if (!pixel_of_rows(A, blockIdx.x, blockIdx.y, x, y)) return;
    int o = y * A.W + x;
    pt_ribbonp r = rib[o];
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ribbon_valid(r)) {
        int sx = clampi((int)((float)x + r.x), 0, A.W - 1), sy = clampi((int)((float)y + r.y), 0, A.H - 1);
        v = A.canvas[sy * A.W + sx];
    }
    A.adv[0].out[o] = v;
}

/* downsampleAverage (downsample.cpp:41-160): one level, CH floats per
 * element, ((a + b) + c) + d then * 0.25 (and * 0.5 for velocities) per
 * channel.  An element's CH floats are loaded together (float2 / float4
 * where CH allows), 64 x 4 threads per block. */
template <int CH> struct DsVec;
template <> struct DsVec<1> { typedef float T; };
template <> struct DsVec<2> { typedef float2 T; };
template <> struct DsVec<4> { typedef float4 T; };

__device__ __forceinline__ float ds_get(const float &v, int) { return v; }
__device__ __forceinline__ float ds_get(const float2 &v, int k) { return k == 0 ? v.x : v.y; }
__device__ __forceinline__ float ds_get(const float4 &v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }
__device__ __forceinline__ void ds_set(float &v, int, float f) { v = f; }
__device__ __forceinline__ void ds_set(float2 &v, int k, float f) { if (k == 0) v.x = f; else v.y = f; }
__device__ __forceinline__ void ds_set(float4 &v, int k, float f)
{
    if (k == 0) v.x = f;
    else if (k == 1) v.y = f;
    else if (k == 2) v.z = f;
    else v.w = f;
}

template <int CH>
__device__ __forceinline__ void b_downsample_v(const float *src, int pw, float *dst, int lw, int lh, int scale_mode,
                                               int bx, int by)
{
    typedef typename DsVec<CH>::T V;
    int x = bx * PT_RX + PT_LX(PT_RX);
    int y = by * PT_RY + PT_LY(PT_RX);
    if (x >= lw || y >= lh) return;
    const V *s = (const V *)src;
    V a = s[(y * 2) * pw + x * 2], b = s[(y * 2) * pw + x * 2 + 1];
    V c = s[(y * 2 + 1) * pw + x * 2], d = s[(y * 2 + 1) * pw + x * 2 + 1];
    V o;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        float v = 0.f;
        v += ds_get(a, k);
        v += ds_get(b, k);
        v += ds_get(c, k);
        v += ds_get(d, k);
        v *= 0.25f;
        if (scale_mode == PT_DOWNSAMPLE_SCALE_WITH_LEVELS)
            v *= 0.5f;
        ds_set(o, k, v);
    }
    ((V *)dst)[y * lw + x] = o;
}

template <int CH>
__global__ __launch_bounds__(PT_RX *PT_RY) void k_downsample_v(const float *src, int pw, float *dst, int lw, int lh,
                                                              int scale_mode)
{
    b_downsample_v<CH>(src, pw, dst, lw, lh, scale_mode, blockIdx.x, blockIdx.y);
}

/* initializeRibbonPyramid (imagePyramid.cpp:211-235): the top-left texel
 * of each 2x2 block as a ribbon -- time step and layer through int, x and y
 * halved -- packed back as (time_step, y, x, layer). */
__global__ __launch_bounds__(PT_RX *PT_RY) void k_downsample_ribbon(const float4 *src, int pw, float4 *dst, int lw,
                                                                   int lh)
{
    int x = blockIdx.x * PT_RX + PT_LX(PT_RX);
    int y = blockIdx.y * PT_RY + PT_LY(PT_RX);
    if (x >= lw || y >= lh) return;
    float4 q = src[(y * 2) * pw + x * 2];
    dst[y * lw + x] = make_float4((float)(int)q.x, q.y * 0.5f, q.z * 0.5f, (float)(int)q.w);
}

/* other element widths (the structure tensor, 3 floats) */
__device__ __forceinline__ void b_downsample(const float *src, int pw, float *dst, int lw, int lh, int ch,
                                             int scale_mode, int bx, int by)
{
    int x = bx * PT_RX + PT_LX(PT_RX);
    int y = by * PT_RY + PT_LY(PT_RX);
    if (x >= lw || y >= lh) return;
    for (int k = 0; k < ch; k++) {
        float s = 0.f;
        s += src[((y * 2) * pw + x * 2) * ch + k];
        s += src[((y * 2) * pw + x * 2 + 1) * ch + k];
        s += src[((y * 2 + 1) * pw + x * 2) * ch + k];
        s += src[((y * 2 + 1) * pw + x * 2 + 1) * ch + k];
        s *= 0.25f;
        if (scale_mode == PT_DOWNSAMPLE_SCALE_WITH_LEVELS)
            s *= 0.5f;
        dst[(y * lw + x) * ch + k] = s;
    }
}

__global__ __launch_bounds__(PT_RX *PT_RY) void k_downsample(const float *src, int pw, float *dst, int lw, int lh,
                                                            int ch, int scale_mode)
{
    b_downsample(src, pw, dst, lw, lh, ch, scale_mode, blockIdx.x, blockIdx.y);
}

/* downsampleVote (downsample.cpp:162-210) for object-ID pyramids: the most
 * frequent of the four IDs, a tie going to the first in scan order. */
__global__ __launch_bounds__(PT_RX *PT_RY) void k_downsample_vote(const int *src, int pw, int *dst, int lw, int lh)
{
    int x = blockIdx.x * PT_RX + PT_LX(PT_RX);
    int y = blockIdx.y * PT_RY + PT_LY(PT_RX);
    if (x >= lw || y >= lh) return;
    int v[4] = {src[(y * 2) * pw + x * 2], src[(y * 2) * pw + x * 2 + 1], src[(y * 2 + 1) * pw + x * 2],
                src[(y * 2 + 1) * pw + x * 2 + 1]};
    int best = v[0], best_n = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int n = (v[0] == v[k]) + (v[1] == v[k]) + (v[2] == v[k]) + (v[3] == v[k]);
        if (n > best_n) {
            best = v[k];
            best_n = n;
        }
    }
    dst[y * lw + x] = best;
}

__global__ void k_fill(float *dst, size_t n, float v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

/* RibbonP(): invalid (layer -1) */
__global__ void k_fill_ribbon(pt_ribbonp *dst, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = pt_ribbonp{0.f, 0.f, 0, -1};
}

/* strucTensorToAngle (imageIO.h:136-139) per element */
__global__ void k_tensor_angle(const float *t, float *a, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = pt_tensor_to_angle(t[3 * i], t[3 * i + 1], t[3 * i + 2]);
}

/* Writes a w x h level into slot `slot` (0 base, 1 output) of the
 * interleaved exemplar texture of width ew. */
__device__ __forceinline__ void b_interleave(const float4 *src, int w, int h, float4 *ex, int ew, int slot, int bx,
                                             int by)
{
    int x = bx * 16 + PT_LX(16);
    int y = by * 16 + PT_LY(16);
    if (x >= w || y >= h) return;
    ex[2 * (y * ew + x) + slot] = src[y * w + x];
}

__global__ void k_interleave(const float4 *src, int w, int h, float4 *ex, int ew, int slot)
{
    b_interleave(src, w, h, ex, ew, slot, blockIdx.x, blockIdx.y);
}

/* createRandomPyramid (synthesisFrame.cpp:102-143) for one level, Philox
 * draws: key (seed, frame*level), counter (x, y, trial, 0), fractional parts
 * from counter (x, y, 50, 0). */
__device__ __forceinline__ void b_random_pyramid(const float4 *color, const float *iorient, int w, int h,
                                                 const float4 *eout, const float *eorient, int ew, int eh, int transp,
                                                 int orient_en, uint32_t key0, uint32_t key1, float4 *out, int bx,
                                                 int by)
{
    int x = bx * 16 + PT_LX(16);
    int y = by * 16 + PT_LY(16);
    if (x >= w || y >= h) return;
    uint32_t key[2] = {key0, key1};
    uint32_t ctr[4] = {(uint32_t)x, (uint32_t)y, 0u, 0u}, o[4];
    int ox, oy, trial = 0;
    float ia = color[y * w + x].w;
    do {
        ctr[2] = (uint32_t)trial;
        pt_philox4x32_10(ctr, key, o);
        ox = pt_rand31(o[0]) % (ew - 1);
        oy = pt_rand31(o[1]) % (eh - 1);
        trial++;
    } while (!transp && ia > 0.99f && eout[oy * ew + ox].w < 0.99f && trial < 50);
    float orientation = orient_en ? iorient[y * w + x] - eorient[oy * ew + ox] : 0.f;
    ctr[2] = 50u;
    pt_philox4x32_10(ctr, key, o);
    float rx = pt_rand_unit(o[0]);
    float ry = pt_rand_unit(o[1]);
    out[y * w + x] = make_float4((float)ox + rx, (float)oy + ry, orientation, 0.0f);
}

__global__ void k_random_pyramid(const float4 *color, const float *iorient, int w, int h, const float4 *eout,
                                 const float *eorient, int ew, int eh, int transp, int orient_en, uint32_t key0,
                                 uint32_t key1, float4 *out)
{
    b_random_pyramid(color, iorient, w, h, eout, eorient, ew, eh, transp, orient_en, key0, key1, out, blockIdx.x,
                     blockIdx.y);
}

/* The same with keyframes: the style of each pixel drawn first from
 * counter (x, y, 51, 0), rand() % keys + first key (synthesisFrame.cpp:
 * 107-110), its exemplar used for the trials and the orientation, the style
 * stored in w. */
struct PtStyleSet {
    const float4 *eout[PT_MAX_STYLE_LAYERS];
    const float *eori[PT_MAX_STYLE_LAYERS];
    int ew[PT_MAX_STYLE_LAYERS], eh[PT_MAX_STYLE_LAYERS];
};

__global__ void k_random_pyramid_keys(const float4 *color, const float *iorient, int w, int h, PtStyleSet S,
                                      int key_first, int num_keys, int transp, int orient_en, uint32_t key0,
                                      uint32_t key1, float4 *out)
{
    int x = blockIdx.x * 16 + PT_LX(16);
    int y = blockIdx.y * 16 + PT_LY(16);
    if (x >= w || y >= h) return;
    uint32_t key[2] = {key0, key1};
    uint32_t ctr[4] = {(uint32_t)x, (uint32_t)y, 51u, 0u}, o[4];
    pt_philox4x32_10(ctr, key, o);
    const int st = pt_rand31(o[0]) % num_keys + key_first;
    const int ew = S.ew[st], eh = S.eh[st];
    const float4 *eout = S.eout[st];
    int ox, oy, trial = 0;
    float ia = color[y * w + x].w;
    do {
        ctr[2] = (uint32_t)trial;
        pt_philox4x32_10(ctr, key, o);
        ox = pt_rand31(o[0]) % (ew - 1);
        oy = pt_rand31(o[1]) % (eh - 1);
        trial++;
    } while (!transp && ia > 0.99f && eout[oy * ew + ox].w < 0.99f && trial < 50);
    float orientation = orient_en ? iorient[y * w + x] - S.eori[st][oy * ew + ox] : 0.f;
    ctr[2] = 50u;
    pt_philox4x32_10(ctr, key, o);
    float rx = pt_rand_unit(o[0]);
    float ry = pt_rand_unit(o[1]);
    out[y * w + x] = make_float4((float)ox + rx, (float)oy + ry, orientation, (float)st);
}

/* PatchXF (36 B, types.h:98-114) <-> SoA planes for the host boundary. */
struct PatchXF36 {
    float x, y, theta, hysteresis;
    int32_t layer;
    float su, sv, ly, la;
};

__device__ __forceinline__ void b_pack_nnf(const PtNnf &s, int W, int lw, int lh, PatchXF36 *d, int bx, int by)
{
    int x = bx * 16 + PT_LX(16);
    int y = by * 16 + PT_LY(16);
    if (x >= lw || y >= lh) return;
    float4 v = s.xyth[y * W + x];
    float2 c = s.sc[y * W + x];
    PatchXF36 p = {v.x, v.y, v.z, v.w, s.la ? s.la[y * W + x] : 0, c.x, c.y, s.lum ? s.lum[y * W + x] : 0.f, 0.f};
    d[y * lw + x] = p;
}

__global__ void k_pack_nnf(PtNnf s, int W, int lw, int lh, PatchXF36 *d) { b_pack_nnf(s, W, lw, lh, d, blockIdx.x, blockIdx.y); }

__global__ void k_unpack_nnf(const PatchXF36 *s, int lw, int lh, PtNnf d, int W)
{
    int x = blockIdx.x * 16 + PT_LX(16);
    int y = blockIdx.y * 16 + PT_LY(16);
    if (x >= lw || y >= lh) return;
    PatchXF36 p = s[y * lw + x];
    d.xyth[y * W + x] = make_float4(p.x, p.y, p.theta, p.hysteresis);
    d.sc[y * W + x] = make_float2(p.su, p.sv);
    if (d.lum) d.lum[y * W + x] = p.ly;
    if (d.la) d.la[y * W + x] = p.layer;
}

/* ------------------------------------------------------------------ */
/* tap-parallel shapes (small levels)                                  */
/* ------------------------------------------------------------------ */
/*
 * On small levels every lane-serial patch distance is a long chain of
 * dependent gathers while most of the GPU idles.  Here each lane computes one
 * tap's residual (the body of calculateResidual's tap loop, identical
 * arithmetic), the values go to LDS, and one lane per candidate then sums
 * them in the reference order (x outer, y inner, sequential fmaf), so the
 * result is the same bit for bit.
 */
template <bool TIME>
__device__ __forceinline__ float tap_residual(const PtArgs &A, __amdgpu_buffer_rsrc_t er, int x, int y, int i, int j,
                                              const Cand &c, const TimeCand &tc)
{
    const int oo = (y + j) * A.W + (x + i);
    const float hw = (float)A.hw, hh = (float)A.hh;
    float4 base = A.color[A.cur][oo];
    float4 canvas = A.canvas[oo];
    const float ibx = (float)i * c.b0x, iby = (float)i * c.b0y;
    float tx = c.t.x + (ibx + (float)j * c.b1x);
    float ty = c.t.y + (iby + (float)j * c.b1y);
    tx = wrap_tap(tx, hw);
    ty = wrap_tap(ty, hh);
    float4 tb, to;
    ex_pair(er, A.ew, A.eh, tx, ty, tb, to);
    float res = ssd_alpha(base, tb, A.alpha_w) * A.in_w;
    if (TIME) {
        bool any_time_w = A.tdi > 0.f || A.tdo > 0.f || A.tcw > 0.f;
        bool fwd_dir = A.direction == PT_FORWARD || A.direction == PT_BIDIRECTIONAL;
        bool bwd_dir = A.direction == PT_BACKWARD || A.direction == PT_BIDIRECTIONAL;
        if (any_time_w) {
            if (fwd_dir && tc.cpv && A.adv[0].xyth[oo].w > 0)
                res = res + time_terms_b(A, er, 0, tc.pf, tc.tb4[0], tc.tb4[1], tc.tb4[2], tc.tb4[3], base, tb,
                                         canvas, to, oo, x, y, i, j);
            if (bwd_dir && tc.cnv && A.adv[1].xyth[oo].w > 0)
                res = res + time_terms_b(A, er, 1, tc.pb, tc.tb4[4], tc.tb4[5], tc.tb4[6], tc.tb4[7], base, tb,
                                         canvas, to, oo, x, y, i, j);
        }
    }
    if (i == 0 && j == 0) {
        if (A.dt_w > 0.f) {
            float de = tex1_lin(A.ex_dt, A.ew, A.ew, A.eh, tx, ty);
            de /= c.sc.y;
            float di = A.dt[A.cur][y * A.W + x];
            float diff = fabsf(de - di);
            diff = fminf(fmaxf(diff, 0.0f), 10.0f);
            res = res + diff * A.dt_w;
        }
    } else {
        res = fmaf(ssd_alpha(canvas, to, A.alpha_w), A.canvas_w, res);
    }
    return res;
}

/* The reference-order sum of one candidate's tap residuals v[t] (t = (i+r)
 * * (2r+1) + (j+r)), starting from acc0. */
template <int RT>
__device__ __forceinline__ float tap_sum(const PtArgs &A, int x, int y, const float *v, float acc0)
{
    constexpr int r = RT, D = 2 * RT + 1;
    float sum_r = acc0, sum_w = 0.f;
    for (int i = -r; i <= r; i++) {
        if (x + i < 0 || x + i >= A.lw) continue;
        for (int j = -r; j <= r; j++) {
            if (y + j < 0 || y + j >= A.lh) continue;
            float wgt = A.gauss[i + r] * A.gauss[j + r];
            sum_r = fmaf(v[(i + r) * D + (j + r)], wgt, sum_r);
            sum_w += wgt;
        }
    }
    if (sum_w < 0.000001f) return 1e31f;
    return sum_r / sum_w;
}

/* Tap-parallel workgroups cover A.tap_np_res (refresh) or A.tap_np_prop
 * (propagate) pixels; the host keeps these within the LDS capacity below
 * (PT_TAP_EVALS tap evaluations per workgroup at most). */
template <int RT> struct TapCfg {
    static constexpr int D = 2 * RT + 1, T = D * D;
    static constexpr int RES_MAX = PT_TAP_EVALS / T;        /* refresh: pixels per workgroup   */
    static constexpr int PROP_MAX = PT_TAP_EVALS / (4 * T); /* propagate: pixels per workgroup */
};

/* per-candidate state of a tap-parallel workgroup; the temporal part only
 * when temporal terms are compiled in */
template <int NU, bool TK> struct TapUnits {
    Cand c[NU];
    int x[NU], y[NU], on[NU];
    TimeCand tc[TK ? NU : 1];
};

template <bool TK, int NU>
__device__ __forceinline__ const TimeCand &tap_tc(const TapUnits<NU, TK> &su, int u)
{
    return su.tc[TK ? u : 0];
}

/* Evaluates taps e = tid, tid + 256, ... of the NU units in `su`: unit e / T,
 * tap e % T (i outer, j inner), into v[e].  Taps outside the level are left
 * unwritten; tap_sum never reads them. */
template <int RT, bool TK, int NUM>
__device__ __forceinline__ void tap_phase(const PtArgs &A, const TapUnits<NUM, TK> &su, int nu, float *v)
{
    constexpr int T = TapCfg<RT>::T, D = 2 * RT + 1;
    __amdgpu_buffer_rsrc_t er = ex_rsrc(A);
    for (int e = threadIdx.x; e < nu * T; e += PT_UP_THREADS) {
        int u = e / T, t = e - u * T;
        if (!su.on[u]) continue;
        int i = t / D - RT, j = t - (t / D) * D - RT;
        int x = su.x[u], y = su.y[u];
        if (x + i < 0 || x + i >= A.lw || y + j < 0 || y + j >= A.lh) continue;
        const TimeCand &tc = tap_tc<TK>(su, u);
        if (TK && (tc.cpv || tc.cnv)) v[e] = tap_residual<true>(A, er, x, y, i, j, su.c[u], tc);
        else v[e] = tap_residual<false>(A, er, x, y, i, j, su.c[u], tc);
    }
}

/* updateResidualCache_kernel, tap-parallel: workgroup g refreshes pixels
 * g*NP .. g*NP+NP-1 of the level in row-major order. */
template <int RT, bool TK> struct ResTapLds {
    TapUnits<TapCfg<RT>::RES_MAX, TK> su;
    float sv[TapCfg<RT>::RES_MAX * TapCfg<RT>::T];
};

template <int RT, bool TK, bool AR = false>
__device__ __forceinline__ void b_update_residual_tp(const PtArgs &A, int g, void *arena = nullptr)
{
    constexpr int T = TapCfg<RT>::T;
    using Lds = ResTapLds<RT, TK>;
    using Units = TapUnits<TapCfg<RT>::RES_MAX, TK>;
    PT_LDS_VAR(Units, su)
    PT_LDS_ARRAY(float, sv, TapCfg<RT>::RES_MAX * T)
    const int tid = threadIdx.x, NP = A.tap_np_res;
    if (tid < NP) {
        int q = g * NP + tid;
        int on = 0;
        if (q < A.lw * A.lh) {
            int x = q % A.lw, y = q / A.lw, o = y * A.W + x;
            su.x[tid] = x;
            su.y[tid] = y;
            if (image_culled(A, x, y)) {
                A.res_base[o] = 0.0f;
            } else {
                on = 1;
                if (A.res_incremental) {
                    on = 0;
                    for (int i = -RT; i <= RT && !on; i++)
                        for (int j = -RT; j <= RT; j++) {
                            int xx = clampi(x + i, 0, A.lw - 1), yy = clampi(y + j, 0, A.lh - 1);
                            if (A.changed[yy * A.W + xx]) {
                                on = 1;
                                break;
                            }
                        }
                }
                if (on) {
                    Cand &c = su.c[tid];
                    c.t = A.base.xyth[o];
                    c.sc = A.base.sc[o];
                    basis(c.t.z, A.su * c.sc.x, A.sv * c.sc.y, c.b0x, c.b0y, c.b1x, c.b1y);
                    if (TK) {
                        TimeCand &tc = su.tc[TK ? tid : 0];
                        tc.acc0 = 0.f;
                        tc.cpv = tc.cnv = false;
                        time_setup<true>(A, x, y, tc.pf, tc.pb, tc.cpv, tc.cnv, tc.tb4);
                        if (tc.cpv) tc.acc0 -= hysteresis_term(A, x, y);
                    }
                }
            }
        }
        su.on[tid] = on;
    }
    __syncthreads();
    tap_phase<RT, TK>(A, su, NP, sv);
    __syncthreads();
    if (tid < NP && su.on[tid]) {
        int x = su.x[tid], y = su.y[tid];
        A.res_base[y * A.W + x] = tap_sum<RT>(A, x, y, sv + tid * T, TK ? tap_tc<TK>(su, tid).acc0 : 0.f);
    }
}

/* propagate_kernel, tap-parallel: workgroup g searches pixels g*NP ..
 * g*NP+NP-1; unit u = 4*pixel + candidate. */
template <int RT, bool TK> struct PropTapLds {
    static constexpr int NUM = 4 * TapCfg<RT>::PROP_MAX;
    TapUnits<NUM, TK> su;
    float sv[NUM * TapCfg<RT>::T];
    float sr[NUM];
};

template <int RT, bool TK, bool AR = false>
__device__ __forceinline__ void b_propagate_tp(const PtArgs &A, int g, int jump, void *arena = nullptr)
{
    constexpr int T = TapCfg<RT>::T, NPM = TapCfg<RT>::PROP_MAX;
    static_assert(NPM >= 1, "tap-parallel propagate needs 4 (2r+1)^2 <= PT_TAP_EVALS");
    using Lds = PropTapLds<RT, TK>;
    using Units = TapUnits<Lds::NUM, TK>;
    PT_LDS_VAR(Units, su)
    PT_LDS_ARRAY(float, sv, Lds::NUM * T)
    PT_LDS_ARRAY(float, sr, Lds::NUM)
    const int tid = threadIdx.x, NP = A.tap_np_prop, NU = 4 * NP;
    if (tid < NU) {
        int k = tid >> 2, cand = tid & 3;
        int q = g * NP + k;
        int on = 0;
        sr[tid] = PT_INF;
        if (q < A.lw * A.lh) {
            int x = q % A.lw, y = q / A.lw, o = y * A.W + x;
            su.x[tid] = x;
            su.y[tid] = y;
            if (!image_culled(A, x, y)) {
                int tx = x, ty = y;
                if (cand == 0) tx = x + jump;
                else if (cand == 1) tx = x - jump;
                else if (cand == 2) ty = y + jump;
                else ty = y - jump;
                Cand &c = su.c[tid];
                propagate_offset(A, tx, ty, (float)x, (float)y, c.t, c.sc);
                if (candidate_testable(A, c.t, A.color[A.cur][o].w)) {
                    on = 1;
                    basis(c.t.z, A.su * c.sc.x, A.sv * c.sc.y, c.b0x, c.b0y, c.b1x, c.b1y);
                    if (TK) {
                        TimeCand &tc = su.tc[TK ? tid : 0];
                        tc.acc0 = 0.f;
                        tc.cpv = tc.cnv = false;
                        time_setup<true>(A, x, y, tc.pf, tc.pb, tc.cpv, tc.cnv, tc.tb4);
                        if (tc.cpv) tc.acc0 -= hysteresis_term(A, x, y);
                    }
                }
            }
        }
        su.on[tid] = on;
    }
    __syncthreads();
    tap_phase<RT, TK>(A, su, NU, sv);
    __syncthreads();
    if (tid < NU && su.on[tid])
        sr[tid] = tap_sum<RT>(A, su.x[tid], su.y[tid], sv + tid * T, TK ? tap_tc<TK>(su, tid).acc0 : 0.f);
    __syncthreads();
    if (tid < NP) {
        int q = g * NP + tid;
        if (q < A.lw * A.lh) {
            int x = q % A.lw, y = q / A.lw, o = y * A.W + x;
            if (image_culled(A, x, y)) {
                A.work.xyth[o] = make_float4(0.f, 0.f, 0.f, 0.f);
                A.work.sc[o] = make_float2(1.f, 1.f);
                lum_set(A, o, false);
            } else {
                float best = A.res_base[o];
                int win = -1;
                for (int k = 0; k < 4; k++) {
                    float v = sr[4 * tid + k];
                    if (v < best) {
                        best = v;
                        win = k;
                    }
                }
                lum_set(A, o, win < 0);
                if (win < 0) {
                    A.work.xyth[o] = A.base.xyth[o];
                    A.work.sc[o] = A.base.sc[o];
                } else {
                    A.work.xyth[o] = su.c[4 * tid + win].t;
                    A.work.sc[o] = su.c[4 * tid + win].sc;
                }
            }
        }
    }
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_UP_THREADS) void k_update_residual_tp(PtArgs A)
{
    b_update_residual_tp<RT, TK>(A, blockIdx.x);
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_UP_THREADS) void k_propagate_tp(PtArgs A, int jump)
{
    b_propagate_tp<RT, TK>(A, blockIdx.x, jump);
}

/* ------------------------------------------------------------------ */
/* multi-job kernels                                                   */
/* ------------------------------------------------------------------ */
/*
 * One launch serves many jobs (frames of different sizes, levels, patch
 * radii and exemplars): the grid is the concatenation of every job's tiles.
 * A workgroup finds its job by binary search over the prefix sums of tile
 * counts (uniform, scalar work), then runs the same body as the single-job
 * kernel on that job's arguments, so results are bit-identical.
 */
struct MultiTile {
    int j, bx, by;
};

/* Kernel parameters of a multi-job launch.  The argument table is read
 * through the constant address space: its loads become scalar loads and the
 * buffer pointers read from it are known to address global memory (global,
 * not flat, memory instructions), as for pointers passed directly as kernel
 * arguments. */
typedef const __attribute__((address_space(4))) PtArgs ConstArgs;
#define PT_MULTI_PARAMS                                                                                   \
    ConstArgs *args, const int *__restrict__ prefix, const int *__restrict__ tiles_x, int njobs, int xcd
#define PT_JOB(j) ((const PtArgs &)args[j])

__device__ __forceinline__ MultiTile multi_tile(const int *__restrict__ prefix, const int *__restrict__ tiles_x,
                                                int njobs, int xcd)
{
    const int n = prefix[njobs];
    int g = blockIdx.x;
    if (xcd) {
        /* workgroups are dealt to the 8 XCDs round-robin: give XCD x the
         * contiguous range of tiles starting at sum_{y<x} count_y so that
         * neighbouring tiles share one L2 */
        int q = n >> 3, rem = n & 7, x = g & 7, k = g >> 3;
        g = x * q + (x < rem ? x : rem) + k;
    }
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    int loc = g - prefix[lo];
    int tx = tiles_x[lo];
    MultiTile t;
    t.j = lo;
    t.bx = loc % tx;
    t.by = loc / tx;
    return t;
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_BX *PT_BY) void km_update_residual(PT_MULTI_PARAMS)
{
    MultiTile t = multi_tile(prefix, tiles_x, njobs, xcd);
    b_update_residual<RT, TK, false>(PT_JOB(t.j), t.bx, t.by);
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_UP_THREADS) void km_update_residual_tp(PT_MULTI_PARAMS)
{
    MultiTile t = multi_tile(prefix, tiles_x, njobs, xcd);
    b_update_residual_tp<RT, TK>(PT_JOB(t.j), t.bx);
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_UP_THREADS) void km_propagate_tp(PT_MULTI_PARAMS)
{
    MultiTile t = multi_tile(prefix, tiles_x, njobs, xcd);
    const PtArgs &A = PT_JOB(t.j);
    b_propagate_tp<RT, TK>(A, t.bx, A.jump);
}

__global__ __launch_bounds__(PT_BX *PT_BY) void km_sync_canvas(PT_MULTI_PARAMS)
{
    MultiTile t = multi_tile(prefix, tiles_x, njobs, xcd);
    const PtArgs &A = PT_JOB(t.j);
    b_sync_canvas(A, t.bx, t.by, A.dw, A.dh);
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_BX *PT_BY) void km_propagate(PT_MULTI_PARAMS)
{
    MultiTile t = multi_tile(prefix, tiles_x, njobs, xcd);
    const PtArgs &A = PT_JOB(t.j);
    b_propagate<RT, TK>(A, t.bx, t.by, A.jump);
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_BX *PT_BY) void km_propagate_ee(PT_MULTI_PARAMS)
{
    MultiTile t = multi_tile(prefix, tiles_x, njobs, xcd);
    const PtArgs &A = PT_JOB(t.j);
    b_propagate_ee<RT, TK>(A, t.bx, t.by, A.jump);
}

template <int RT, bool EE, int TX, int TY, bool TM, bool XL = false>
__global__ __launch_bounds__(PT_UP_THREADS) __attribute__((amdgpu_waves_per_eu(TM ? 4 : 1))) void
km_propagate_up(PT_MULTI_PARAMS)
{
    MultiTile t = multi_tile(prefix, tiles_x, njobs, xcd);
    const PtArgs &A = PT_JOB(t.j);
    b_propagate_up<RT, EE, TX, TY, TM, XL>(A, t.bx, t.by, A.jump);
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_BX *PT_BY) void km_scatter(PT_MULTI_PARAMS)
{
    MultiTile t = multi_tile(prefix, tiles_x, njobs, xcd);
    b_scatter<RT, TK>(PT_JOB(t.j), t.bx, t.by);
}

template <int RT, bool TK>
__global__ __launch_bounds__(PT_BX *PT_BY) void km_scatter_ee(PT_MULTI_PARAMS)
{
    MultiTile t = multi_tile(prefix, tiles_x, njobs, xcd);
    b_scatter_ee<RT, TK>(PT_JOB(t.j), t.bx, t.by);
}

template <int RT, bool EE, int TX, int TY, bool TM>
__global__ __launch_bounds__(PT_UP_THREADS) void km_scatter_up(PT_MULTI_PARAMS)
{
    MultiTile t = multi_tile(prefix, tiles_x, njobs, xcd);
    const PtArgs &A = PT_JOB(t.j);
    b_scatter_up<RT, EE, TX, TY, TM>(A, t.bx, t.by, A.nscat);
}

/* ------------------------------------------------------------------ */
/* launchers                                                           */
/* ------------------------------------------------------------------ */
/* Task capture (pt_sink_begin): while a sink is active on the calling
 * thread, a launcher appends the task form of its kernel instead of
 * launching it. */
static thread_local PtTaskSink *g_sink = nullptr;

/* elements per workgroup tile of the copy and fill tasks (4 per lane) */
#define PT_TASK_ELEMS 1024

void pt_sink_begin(PtTaskSink *s) { g_sink = s; }
void pt_sink_end() { g_sink = nullptr; }

static PtTask task_new(int kind, int64_t ntiles, int tiles_x)
{
    PtTask t;
    memset(&t, 0, sizeof(t));
    t.kind = kind;
    t.ntiles = (int)ntiles;
    t.tiles_x = tiles_x > 0 ? tiles_x : 1;
    t.arg = -1;
    return t;
}

/* tasks without tiles are dropped: they do no work */
static hipError_t sink_push(const PtTask &t, const PtArgs *a = nullptr)
{
    if (t.ntiles <= 0) return hipSuccess;
    PtTask u = t;
    if (a) {
        u.arg = (int)g_sink->args.size();
        g_sink->args.push_back(*a);
    }
    g_sink->tasks.push_back(u);
    return hipSuccess;
}

static hipError_t sink_refuse(const char *what)
{
    if (!g_sink->unsupported) g_sink->unsupported = what;
    return hipErrorNotSupported;
}

/* a 2-D tiling of w x h pixels in tx x ty tiles */
static PtTask task_2d(int kind, int w, int h, int tx, int ty)
{
    int nx = w > 0 ? (w + tx - 1) / tx : 0, ny = h > 0 ? (h + ty - 1) / ty : 0;
    return task_new(kind, (int64_t)nx * ny, nx);
}

#define PT_SINK_REFUSE(name)                                                                              \
    do {                                                                                                  \
        if (g_sink) return sink_refuse(name);                                                             \
    } while (0)

/* the extended patch-distance terms (coherence, LAB, object IDs) and style
 * layers (keyframe interpolation) run in the per-pixel kernels' RT = -1
 * variant */
static bool ext_terms(const PtArgs &a) { return a.coh_w > 0.f || a.lab || a.obj_w > 0.f || a.layered; }

/* ... and, at radius 2 or 3 without style layers, the default unit-parallel
 * propagate and scatter shapes' RT = -2, -3 variants */
static bool ext_up(const PtArgs &a) { return ext_terms(a) && !a.layered && !a.keys && (a.r == 2 || a.r == 3); }

static dim3 grid_for(int w, int h) { return dim3((w + PT_BX - 1) / PT_BX, (h + PT_BY - 1) / PT_BY); }
static const dim3 kBlock(PT_BX * PT_BY);
static dim3 grid_rows(int w, int h) { return dim3((w + PT_RX - 1) / PT_RX, (h + PT_RY - 1) / PT_RY); }
static const dim3 kRowBlock(PT_RX * PT_RY);

#define PT_DISPATCH_R(KERNEL, A, ...)                                                                     \
    do {                                                                                                  \
        dim3 g = grid_for((A).lw, (A).lh);                                                                \
        if ((A).time_active) {                                                                            \
            switch ((A).r) {                                                                              \
            case 2: hipLaunchKernelGGL((KERNEL<2, true>), g, kBlock, 0, s, (A), ##__VA_ARGS__); break;    \
            case 3: hipLaunchKernelGGL((KERNEL<3, true>), g, kBlock, 0, s, (A), ##__VA_ARGS__); break;    \
            case 4: hipLaunchKernelGGL((KERNEL<4, true>), g, kBlock, 0, s, (A), ##__VA_ARGS__); break;    \
            default: hipLaunchKernelGGL((KERNEL<0, true>), g, kBlock, 0, s, (A), ##__VA_ARGS__); break;   \
            }                                                                                             \
        } else {                                                                                          \
            switch ((A).r) {                                                                              \
            case 2: hipLaunchKernelGGL((KERNEL<2, false>), g, kBlock, 0, s, (A), ##__VA_ARGS__); break;   \
            case 3: hipLaunchKernelGGL((KERNEL<3, false>), g, kBlock, 0, s, (A), ##__VA_ARGS__); break;   \
            case 4: hipLaunchKernelGGL((KERNEL<4, false>), g, kBlock, 0, s, (A), ##__VA_ARGS__); break;   \
            default: hipLaunchKernelGGL((KERNEL<0, false>), g, kBlock, 0, s, (A), ##__VA_ARGS__); break;  \
            }                                                                                             \
        }                                                                                                 \
    } while (0)

static hipError_t sink_op(int op, const PtArgs &a)
{
    PtTask t;
    const char *why = nullptr;
    if (pt_task_op(op, pt_default_shape(op, a), a, 0, &t, &why) != PT_OK) return sink_refuse(why);
    return sink_push(t, &a);
}

hipError_t pt_launch_update_residual(const PtArgs &a, hipStream_t s)
{
    if (g_sink) return sink_op(PT_OP_RESIDUAL, a);
    if (pt_default_shape(PT_OP_RESIDUAL, a) == PT_SHAPE_TAP) {
        dim3 g(pt_shape_tiles(PT_OP_RESIDUAL, PT_SHAPE_TAP, a, nullptr)), bu(PT_UP_THREADS);
        bool tk = a.time_active != 0;
        switch (a.r) {
        case 2: hipLaunchKernelGGL((tk ? k_update_residual_tp<2, true> : k_update_residual_tp<2, false>), g, bu, 0, s, a); break;
        case 3: hipLaunchKernelGGL((tk ? k_update_residual_tp<3, true> : k_update_residual_tp<3, false>), g, bu, 0, s, a); break;
        default: hipLaunchKernelGGL((tk ? k_update_residual_tp<4, true> : k_update_residual_tp<4, false>), g, bu, 0, s, a); break;
        }
        return hipGetLastError();
    }
    bool tk = a.time_active != 0;
    if (ext_terms(a)) {
        hipLaunchKernelGGL((tk ? k_update_residual<-1, true> : k_update_residual<-1, false>), grid_for(a.lw, a.lh),
                           kBlock, 0, s, a);
        return hipGetLastError();
    }
    if (a.ex_window && a.ew >= PT_XW_RES && a.eh >= PT_XW_RES && (a.r == 2 || a.r == 3)) {
        dim3 g = grid_for(a.lw, a.lh);
        if (a.r == 2)
            hipLaunchKernelGGL((tk ? k_update_residual<2, true, true> : k_update_residual<2, false, true>), g, kBlock,
                               0, s, a);
        else
            hipLaunchKernelGGL((tk ? k_update_residual<3, true, true> : k_update_residual<3, false, true>), g, kBlock,
                               0, s, a);
        return hipGetLastError();
    }
    PT_DISPATCH_R(k_update_residual, a);
    return hipGetLastError();
}

hipError_t pt_launch_canvas(const PtArgs &a, hipStream_t s)
{
    if (g_sink) {
        PtArgs c = a;
        c.dw = c.dh = 0;
        return sink_op(PT_OP_SYNC_CANVAS, c);
    }
    hipLaunchKernelGGL(k_canvas, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a);
    return hipGetLastError();
}

hipError_t pt_launch_sync_canvas(const PtArgs &a, int dw, int dh, hipStream_t s)
{
    if (g_sink) {
        PtArgs c = a;
        c.dw = dw;
        c.dh = dh;
        return sink_op(PT_OP_SYNC_CANVAS, c);
    }
    int gw = dw > a.lw ? dw : a.lw, gh = dh > a.lh ? dh : a.lh;
    hipLaunchKernelGGL(k_sync_canvas, grid_rows(gw, gh), kRowBlock, 0, s, a, dw, dh);
    return hipGetLastError();
}

hipError_t pt_launch_copy_random(const PtArgs &a, hipStream_t s)
{
    if (g_sink) return sink_push(task_2d(PT_TASK_COPY_RANDOM, a.lw, a.lh, PT_RX, PT_RY), &a);
    hipLaunchKernelGGL(k_copy_random, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a);
    return hipGetLastError();
}

#define PT_UP_LAUNCH(KERNEL, EE, TX, TY, A, ...)                                                         \
    do {                                                                                                  \
        dim3 g(((A).lw + (TX)-1) / (TX), ((A).lh + (TY)-1) / (TY));                                       \
        dim3 b(PT_UP_THREADS);                                                                            \
        switch ((A).r) {                                                                                  \
        case 2: hipLaunchKernelGGL((KERNEL<2, EE, TX, TY>), g, b, 0, s, (A), ##__VA_ARGS__); break;       \
        case 3: hipLaunchKernelGGL((KERNEL<3, EE, TX, TY>), g, b, 0, s, (A), ##__VA_ARGS__); break;       \
        case 4: hipLaunchKernelGGL((KERNEL<4, EE, TX, TY>), g, b, 0, s, (A), ##__VA_ARGS__); break;       \
        default: hipLaunchKernelGGL((KERNEL<0, EE, TX, TY>), g, b, 0, s, (A), ##__VA_ARGS__); break;      \
        }                                                                                                 \
    } while (0)

/* samples of the reference scatter loop (texSynth_kernel.cu:909): identical
 * for every pixel of a launch */
static int scatter_count(const PtArgs &a)
{
    int n = 0;
    float dist = (float)(a.ew > a.eh ? a.ew : a.eh);
    for (int i = 0; (i < a.samples) && (dist * 0.9f >= 10.f); i++) {
        n++;
        dist *= 0.9f;
    }
    return n;
}

/* pixel tile of (op, shape) */
static void shape_tile(int op, int shape, int &tx, int &ty)
{
    tx = PT_BX;
    ty = PT_BY;
    if (op == PT_OP_SYNC_CANVAS) {
        tx = PT_RX;
        ty = PT_RY;
        return;
    }
    if (op == PT_OP_PROPAGATE) {
        if (shape == PT_SHAPE_UP_EE_A || shape == PT_SHAPE_UP_B) { tx = 16; ty = 8; }
        else if (shape == PT_SHAPE_UP_A || shape == PT_SHAPE_UP_EE_B) tx = ty = 8;
    } else if (op == PT_OP_SCATTER) {
        if (shape == PT_SHAPE_UP_EE_A) { tx = 16; ty = 8; }
        else if (shape == PT_SHAPE_UP_A) { tx = 4; ty = 4; }
        else if (shape == PT_SHAPE_UP_EE_B || shape == PT_SHAPE_UP_B) { tx = 8; ty = 8; }
    }
}

static bool shape_is_up(int shape) { return shape >= PT_SHAPE_UP_EE_A; }
static bool shape_is_ee(int shape)
{
    return shape == PT_SHAPE_PIXEL_EE || shape == PT_SHAPE_UP_EE_A || shape == PT_SHAPE_UP_EE_B;
}

/* pixels per tap-parallel workgroup for (op, a.r); 0 when unsupported */
static int tap_pixels(int op, const PtArgs &a)
{
    if (op == PT_OP_RESIDUAL) return (a.r >= 2 && a.r <= 4) ? a.tap_np_res : 0;
    if (op == PT_OP_PROPAGATE) return (a.r >= 2 && a.r <= 3) ? a.tap_np_prop : 0;
    return 0;
}

void pt_tap_configure(PtArgs &a, int evals)
{
    int t = (2 * a.r + 1) * (2 * a.r + 1);
    evals = evals < 1 ? 1 : (evals > PT_TAP_EVALS ? PT_TAP_EVALS : evals);
    int nr = evals / t, np = evals / (4 * t);
    a.tap_np_res = nr < 1 ? 1 : nr;
    a.tap_np_prop = np < 1 ? 1 : np;
}

int pt_shape_valid(int op, int shape, const PtArgs &a)
{
    if (shape < 0 || shape >= PT_SHAPE_COUNT) return 0;
    /* the coherence term, the LAB colour space, the object-ID term and style
     * layers are evaluated by the per-pixel in-order kernels (their RT = -1
     * variant) and, where ext_up holds, by the unit-parallel shapes that
     * pt_default_shape picks for propagate (8x8, with and without early
     * rejection) and scatter (16x8 with, 4x4 without) */
    if (ext_terms(a) && shape != PT_SHAPE_PIXEL) {
        if (!ext_up(a)) return 0;
        if (op == PT_OP_PROPAGATE) {
            if (shape != PT_SHAPE_UP_EE_B && shape != PT_SHAPE_UP_A) return 0;
        } else if (op == PT_OP_SCATTER) {
            if (shape != PT_SHAPE_UP_EE_A && shape != PT_SHAPE_UP_A) return 0;
        } else {
            return 0;
        }
    }
    if (shape == PT_SHAPE_TAP) return tap_pixels(op, a) > 0;
    if (op == PT_OP_RESIDUAL || op == PT_OP_SYNC_CANVAS) return shape == PT_SHAPE_PIXEL;
    if (op != PT_OP_PROPAGATE && op != PT_OP_SCATTER) return 0;
    if (shape_is_ee(shape) && !a.ee_ok) return 0;
    if (shape_is_up(shape)) {
        if (op == PT_OP_SCATTER && a.nscat > PT_UP_MAX_SAMPLES) return 0;
    }
    return 1;
}

int pt_default_shape(int op, const PtArgs &a)
{
    if ((int64_t)a.lw * a.lh <= a.tap_max_px && pt_shape_valid(op, PT_SHAPE_TAP, a)) return PT_SHAPE_TAP;
    if (op != PT_OP_PROPAGATE && op != PT_OP_SCATTER) return PT_SHAPE_PIXEL;
    if (a.unit_parallel && pt_shape_valid(op, PT_SHAPE_UP_A, a)) {
        /* propagate keeps a candidate table per unit in LDS: the 8x8 tile
         * keeps occupancy up */
        if (a.early_exit) return op == PT_OP_PROPAGATE ? PT_SHAPE_UP_EE_B : PT_SHAPE_UP_EE_A;
        return PT_SHAPE_UP_A;
    }
    return (a.early_exit && pt_shape_valid(op, PT_SHAPE_PIXEL_EE, a)) ? PT_SHAPE_PIXEL_EE : PT_SHAPE_PIXEL;
}

int pt_shape_tiles(int op, int shape, const PtArgs &a, int *tiles_x)
{
    if (shape == PT_SHAPE_TAP) {
        int np = tap_pixels(op, a);
        int n = np > 0 ? (a.lw * a.lh + np - 1) / np : 0;
        if (tiles_x) *tiles_x = n > 0 ? n : 1;
        return n;
    }
    int tx, ty;
    shape_tile(op, shape, tx, ty);
    int w = a.lw, h = a.lh;
    if (op == PT_OP_SYNC_CANVAS) {
        w = a.dw > w ? a.dw : w;
        h = a.dh > h ? a.dh : h;
    }
    int nx = (w + tx - 1) / tx, ny = (h + ty - 1) / ty;
    if (tiles_x) *tiles_x = nx;
    return nx * ny;
}

static int rt_of(int r) { return (r == 2 || r == 3 || r == 4) ? r : 0; }

/* Whether a propagate launch of this shape takes the exemplar-window
 * variant: a short jump (a.jump) on a large level, a fixed radius, an
 * exemplar that holds the window, an 8x8 unit-parallel shape. */
static bool prop_xl(int shape, const PtArgs &a)
{
    return a.prop_window && a.jump <= a.prop_window_jump && a.ew >= PT_XW_PROP && a.eh >= PT_XW_PROP &&
           (a.r == 2 || a.r == 3) && (shape == PT_SHAPE_UP_EE_B || shape == PT_SHAPE_UP_A) && !ext_terms(a);
}

int pt_shape_key(int op, int shape, const PtArgs &a)
{
    int tk = (op == PT_OP_SYNC_CANVAS) ? 0 : (a.time_active ? 1 : 0);
    int rt = (op == PT_OP_SYNC_CANVAS) ? 0 : (ext_terms(a) ? (shape_is_up(shape) ? 4 + a.r : 1) : rt_of(a.r));
    int xl = (op == PT_OP_PROPAGATE && prop_xl(shape, a)) ? 1 : 0;
    return (((op * PT_SHAPE_COUNT + shape) * 8 + rt) * 2 + tk) * 2 + xl;
}

/* One kernel instantiation per (op, shape, RT, TK); `L` launches it either
 * as a single-job kernel or as a multi-job kernel. */
#define PT_RT_SWITCH(r, F)                                                                                \
    switch (rt_of(r)) {                                                                                   \
    case 2: F(2); break;                                                                                  \
    case 3: F(3); break;                                                                                  \
    case 4: F(4); break;                                                                                  \
    default: F(0); break;                                                                                 \
    }
#define PT_RTX_SWITCH(a, F)                                                                               \
    if (ext_terms(a)) {                                                                                   \
        F(-1);                                                                                            \
    } else {                                                                                              \
        PT_RT_SWITCH((a).r, F)                                                                            \
    }
/* the unit-parallel shapes that carry the extended terms (ext_up) */
#define PT_RTU_SWITCH(a, F)                                                                               \
    if (ext_terms(a)) {                                                                                   \
        if ((a).r == 2) F(-2);                                                                            \
        else F(-3);                                                                                       \
    } else {                                                                                              \
        PT_RT_SWITCH((a).r, F)                                                                            \
    }

hipError_t pt_launch_propagate(const PtArgs &a0, int jump, hipStream_t s)
{
    PtArgs a = a0;
    a.jump = jump;
    if (g_sink) return sink_op(PT_OP_PROPAGATE, a);
    int shape = pt_default_shape(PT_OP_PROPAGATE, a);
    bool tk = a.time_active != 0;
    if (shape == PT_SHAPE_TAP) {
        dim3 gt(pt_shape_tiles(PT_OP_PROPAGATE, PT_SHAPE_TAP, a, nullptr)), bt(PT_UP_THREADS);
        if (a.r == 2) hipLaunchKernelGGL((tk ? k_propagate_tp<2, true> : k_propagate_tp<2, false>), gt, bt, 0, s, a, jump);
        else hipLaunchKernelGGL((tk ? k_propagate_tp<3, true> : k_propagate_tp<3, false>), gt, bt, 0, s, a, jump);
        return hipGetLastError();
    }
    int tx, ty;
    shape_tile(PT_OP_PROPAGATE, shape, tx, ty);
    dim3 g((a.lw + tx - 1) / tx, (a.lh + ty - 1) / ty), bu(PT_UP_THREADS);
#define L_PIX(R) hipLaunchKernelGGL((tk ? k_propagate<R, true> : k_propagate<R, false>), g, kBlock, 0, s, a, jump)
#define L_PEE(R) hipLaunchKernelGGL((tk ? k_propagate_ee<R, true> : k_propagate_ee<R, false>), g, kBlock, 0, s, a, jump)
#define L_UEB(R) hipLaunchKernelGGL((tk ? k_propagate_up<R, true, 8, 8, true> : k_propagate_up<R, true, 8, 8, false>), g, bu, 0, s, a, jump)
#define L_UA(R) hipLaunchKernelGGL((tk ? k_propagate_up<R, false, 8, 8, true> : k_propagate_up<R, false, 8, 8, false>), g, bu, 0, s, a, jump)
    /* exemplar windows for the short jumps on large levels, whose candidates
     * lie near the tile's current offsets */
    if (prop_xl(shape, a)) {
        const bool ee = shape == PT_SHAPE_UP_EE_B;
#define L_XL(R)                                                                                           \
    hipLaunchKernelGGL((ee ? (tk ? k_propagate_up<R, true, 8, 8, true, true> : k_propagate_up<R, true, 8, 8, false, true>) \
                           : (tk ? k_propagate_up<R, false, 8, 8, true, true> : k_propagate_up<R, false, 8, 8, false, true>)), \
                       g, bu, 0, s, a, jump)
        if (a.r == 2) L_XL(2);
        else L_XL(3);
#undef L_XL
        return hipGetLastError();
    }
    switch (shape) {
    case PT_SHAPE_PIXEL: PT_RTX_SWITCH(a, L_PIX); break;
    case PT_SHAPE_PIXEL_EE: PT_RT_SWITCH(a.r, L_PEE); break;
    case PT_SHAPE_UP_EE_B: PT_RTU_SWITCH(a, L_UEB); break;
    default: PT_RTU_SWITCH(a, L_UA); break;
    }
#undef L_PIX
#undef L_PEE
#undef L_UEB
#undef L_UA
    return hipGetLastError();
}

hipError_t pt_launch_scatter(const PtArgs &a0, hipStream_t s)
{
    PtArgs a = a0;
    a.nscat = scatter_count(a);
    if (g_sink) return sink_op(PT_OP_SCATTER, a);
    const int n = a.nscat;
    int shape = pt_default_shape(PT_OP_SCATTER, a);
    int tx, ty;
    shape_tile(PT_OP_SCATTER, shape, tx, ty);
    dim3 g((a.lw + tx - 1) / tx, (a.lh + ty - 1) / ty), bu(PT_UP_THREADS);
    bool tk = a.time_active != 0;
#define L_PIX(R) hipLaunchKernelGGL((tk ? k_scatter<R, true> : k_scatter<R, false>), g, kBlock, 0, s, a)
#define L_PEE(R) hipLaunchKernelGGL((tk ? k_scatter_ee<R, true> : k_scatter_ee<R, false>), g, kBlock, 0, s, a)
#define L_UEA(R) hipLaunchKernelGGL((tk ? k_scatter_up<R, true, 16, 8, true> : k_scatter_up<R, true, 16, 8, false>), g, bu, 0, s, a, n)
#define L_UA(R) hipLaunchKernelGGL((tk ? k_scatter_up<R, false, 4, 4, true> : k_scatter_up<R, false, 4, 4, false>), g, bu, 0, s, a, n)
    switch (shape) {
    case PT_SHAPE_PIXEL: PT_RTX_SWITCH(a, L_PIX); break;
    case PT_SHAPE_PIXEL_EE: PT_RT_SWITCH(a.r, L_PEE); break;
    case PT_SHAPE_UP_EE_A: PT_RTU_SWITCH(a, L_UEA); break;
    default: PT_RTU_SWITCH(a, L_UA); break;
    }
#undef L_PIX
#undef L_PEE
#undef L_UEA
#undef L_UA
    return hipGetLastError();
}

hipError_t pt_launch_multi(int op, int shape, const PtArgs &a, const PtJobTable &t, int ntiles, hipStream_t s)
{
    PT_SINK_REFUSE("multi-job launch");
    if (ntiles <= 0) return hipSuccess;
    if (!pt_shape_valid(op, shape, a)) return hipErrorInvalidValue;
    dim3 g(ntiles), bu(PT_UP_THREADS);
    bool tk = a.time_active != 0;
    switch (op) {
    case PT_OP_RESIDUAL: {
        if (shape == PT_SHAPE_TAP) {
#define LT(R) hipLaunchKernelGGL((tk ? km_update_residual_tp<R, true> : km_update_residual_tp<R, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
            if (a.r == 2) LT(2);
            else if (a.r == 3) LT(3);
            else LT(4);
#undef LT
            break;
        }
#define L(R) hipLaunchKernelGGL((tk ? km_update_residual<R, true> : km_update_residual<R, false>), g, kBlock, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
        PT_RTX_SWITCH(a, L);
#undef L
        break;
    }
    case PT_OP_SYNC_CANVAS: hipLaunchKernelGGL(km_sync_canvas, g, kRowBlock, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap); break;
    case PT_OP_PROPAGATE: {
#define L_PIX(R) hipLaunchKernelGGL((tk ? km_propagate<R, true> : km_propagate<R, false>), g, kBlock, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
#define L_PEE(R) hipLaunchKernelGGL((tk ? km_propagate_ee<R, true> : km_propagate_ee<R, false>), g, kBlock, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
#define L_UEA(R) hipLaunchKernelGGL((tk ? km_propagate_up<R, true, 16, 8, true> : km_propagate_up<R, true, 16, 8, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
#define L_UA(R) hipLaunchKernelGGL((tk ? km_propagate_up<R, false, 8, 8, true> : km_propagate_up<R, false, 8, 8, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
#define L_UEB(R) hipLaunchKernelGGL((tk ? km_propagate_up<R, true, 8, 8, true> : km_propagate_up<R, true, 8, 8, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
#define L_UB(R) hipLaunchKernelGGL((tk ? km_propagate_up<R, false, 16, 8, true> : km_propagate_up<R, false, 16, 8, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
        if (prop_xl(shape, a)) {
            const bool ee = shape == PT_SHAPE_UP_EE_B;
#define L_XL(R)                                                                                           \
    hipLaunchKernelGGL((ee ? (tk ? km_propagate_up<R, true, 8, 8, true, true> : km_propagate_up<R, true, 8, 8, false, true>) \
                           : (tk ? km_propagate_up<R, false, 8, 8, true, true> : km_propagate_up<R, false, 8, 8, false, true>)), \
                       g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
            if (a.r == 2) L_XL(2);
            else L_XL(3);
#undef L_XL
            break;
        }
        switch (shape) {
        case PT_SHAPE_PIXEL: PT_RTX_SWITCH(a, L_PIX); break;
        case PT_SHAPE_PIXEL_EE: PT_RT_SWITCH(a.r, L_PEE); break;
        case PT_SHAPE_UP_EE_A: PT_RT_SWITCH(a.r, L_UEA); break;
        case PT_SHAPE_UP_A: PT_RTU_SWITCH(a, L_UA); break;
        case PT_SHAPE_UP_EE_B: PT_RTU_SWITCH(a, L_UEB); break;
        case PT_SHAPE_TAP:
            if (a.r == 2) hipLaunchKernelGGL((tk ? km_propagate_tp<2, true> : km_propagate_tp<2, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap);
            else hipLaunchKernelGGL((tk ? km_propagate_tp<3, true> : km_propagate_tp<3, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap);
            break;
        default: PT_RT_SWITCH(a.r, L_UB); break;
        }
#undef L_PIX
#undef L_PEE
#undef L_UEA
#undef L_UA
#undef L_UEB
#undef L_UB
        break;
    }
    case PT_OP_SCATTER: {
#define L_PIX(R) hipLaunchKernelGGL((tk ? km_scatter<R, true> : km_scatter<R, false>), g, kBlock, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
#define L_PEE(R) hipLaunchKernelGGL((tk ? km_scatter_ee<R, true> : km_scatter_ee<R, false>), g, kBlock, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
#define L_UEA(R) hipLaunchKernelGGL((tk ? km_scatter_up<R, true, 16, 8, true> : km_scatter_up<R, true, 16, 8, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
#define L_UA(R) hipLaunchKernelGGL((tk ? km_scatter_up<R, false, 4, 4, true> : km_scatter_up<R, false, 4, 4, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
#define L_UEB(R) hipLaunchKernelGGL((tk ? km_scatter_up<R, true, 8, 8, true> : km_scatter_up<R, true, 8, 8, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
#define L_UB(R) hipLaunchKernelGGL((tk ? km_scatter_up<R, false, 8, 8, true> : km_scatter_up<R, false, 8, 8, false>), g, bu, 0, s, (ConstArgs *)t.args, t.prefix, t.tiles_x, t.n, t.xcd_remap)
        switch (shape) {
        case PT_SHAPE_PIXEL: PT_RTX_SWITCH(a, L_PIX); break;
        case PT_SHAPE_PIXEL_EE: PT_RT_SWITCH(a.r, L_PEE); break;
        case PT_SHAPE_UP_EE_A: PT_RTU_SWITCH(a, L_UEA); break;
        case PT_SHAPE_UP_A: PT_RTU_SWITCH(a, L_UA); break;
        case PT_SHAPE_UP_EE_B: PT_RT_SWITCH(a.r, L_UEB); break;
        default: PT_RT_SWITCH(a.r, L_UB); break;
        }
#undef L_PIX
#undef L_PEE
#undef L_UEA
#undef L_UA
#undef L_UEB
#undef L_UB
        break;
    }
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t pt_launch_upsample(const PtArgs &a, hipStream_t s)
{
    if (g_sink) return sink_push(task_2d(PT_TASK_UPSAMPLE, a.lw, a.lh, PT_RX, PT_RY), &a);
    hipLaunchKernelGGL(k_upsample, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a);
    return hipGetLastError();
}

hipError_t pt_launch_advect(const PtArgs &a, int forwards, int time_step, int flags, hipStream_t s)
{
    PT_SINK_REFUSE("advection");
    hipLaunchKernelGGL(k_advect, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a, forwards, time_step, flags);
    return hipGetLastError();
}

hipError_t pt_launch_time_table(const PtArgs &a, int k, float4 *tt, hipStream_t s)
{
    PT_SINK_REFUSE("time table");
    hipLaunchKernelGGL(k_time_table, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a, k, tt);
    return hipGetLastError();
}

hipError_t pt_launch_merge_randomly(const PtArgs &a, int forwards, float weight, int force, hipStream_t s)
{
    PT_SINK_REFUSE("random merge");
    hipLaunchKernelGGL(k_merge_randomly, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a, forwards, weight, force);
    return hipGetLastError();
}

hipError_t pt_launch_merge_residual(const PtArgs &a, int forwards, hipStream_t s)
{
    PT_SINK_REFUSE("residual merge");
    hipLaunchKernelGGL(k_merge_residual, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a, forwards);
    return hipGetLastError();
}

hipError_t pt_launch_merge_advected(const PtArgs &a, int fallback, hipStream_t s)
{
    PT_SINK_REFUSE("advected merge");
    hipLaunchKernelGGL(k_merge_advected, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a, fallback);
    return hipGetLastError();
}

hipError_t pt_launch_offsets_from_ribbon(const PtArgs &a, int look_forwards, hipStream_t s)
{
    PT_SINK_REFUSE("ribbon offsets");
    hipLaunchKernelGGL(k_offsets_from_ribbon, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a, look_forwards);
    return hipGetLastError();
}

hipError_t pt_launch_accumulate_ribbon(const PtArgs &a, const pt_ribbonp *src, pt_ribbonp *dst, int forwards,
                                       int time_step, int overwrite, hipStream_t s)
{
    PT_SINK_REFUSE("ribbon accumulation");
    hipLaunchKernelGGL(k_accumulate_ribbon, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a, src, dst, forwards, time_step,
                       overwrite);
    return hipGetLastError();
}

hipError_t pt_launch_ribbon_output(const PtArgs &a, const pt_ribbonp *rib, hipStream_t s)
{
    PT_SINK_REFUSE("ribbon output");
    hipLaunchKernelGGL(k_ribbon_output, grid_rows(a.lw, a.lh), kRowBlock, 0, s, a, rib);
    return hipGetLastError();
}

hipError_t pt_launch_downsample_ribbon(const float *src, int pw, float *dst, int lw, int lh, hipStream_t s)
{
    PT_SINK_REFUSE("ribbon pyramid");
    hipLaunchKernelGGL(k_downsample_ribbon, grid_rows(lw, lh), kRowBlock, 0, s, (const float4 *)src, pw,
                       (float4 *)dst, lw, lh);
    return hipGetLastError();
}

hipError_t pt_launch_fill_ribbon(pt_ribbonp *dst, size_t n, hipStream_t s)
{
    PT_SINK_REFUSE("ribbon fill");
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_ribbon, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, n);
    return hipGetLastError();
}

hipError_t pt_launch_random_pyramid_keys(const float4 *color, const float *iorient, int w, int h, int nstyles,
                                         const float4 *const *eouts, const float *const *eorients, const int *ews,
                                         const int *ehs, int key_first, int num_keys, int transp, int orient_en,
                                         uint32_t key0, uint32_t key1, float4 *out, hipStream_t s)
{
    PT_SINK_REFUSE("keyed random pyramid");
    if (nstyles < 1 || nstyles > PT_MAX_STYLE_LAYERS || num_keys < 1 || key_first < 0 ||
        key_first + num_keys > nstyles)
        return hipErrorInvalidValue;
    PtStyleSet S{};
    for (int k = 0; k < nstyles; k++) {
        S.eout[k] = eouts[k];
        S.eori[k] = eorients[k];
        S.ew[k] = ews[k];
        S.eh[k] = ehs[k];
    }
    dim3 b(256), g((w + 15) / 16, (h + 15) / 16);
    hipLaunchKernelGGL(k_random_pyramid_keys, g, b, 0, s, color, iorient, w, h, S, key_first, num_keys, transp,
                       orient_en, key0, key1, out);
    return hipGetLastError();
}

hipError_t pt_launch_downsample(const float *src, int pw, float *dst, int lw, int lh, int ch, int scale_mode,
                                hipStream_t s)
{
    if (g_sink) {
        if (ch < 1 || ch > 4 || (scale_mode != PT_DOWNSAMPLE_SCALE_CONSTANT && scale_mode != PT_DOWNSAMPLE_SCALE_WITH_LEVELS))
            return sink_refuse("downsample mode");
        PtTask t = task_2d(PT_TASK_DOWNSAMPLE, lw, lh, PT_RX, PT_RY);
        t.variant = ch;
        t.p[0] = src;
        t.q = dst;
        t.i[0] = pw;
        t.i[1] = lw;
        t.i[2] = lh;
        t.i[3] = scale_mode;
        return sink_push(t);
    }
    dim3 b = kRowBlock, g = grid_rows(lw, lh);
    switch (ch) {
    case 1: hipLaunchKernelGGL(k_downsample_v<1>, g, b, 0, s, src, pw, dst, lw, lh, scale_mode); break;
    case 2: hipLaunchKernelGGL(k_downsample_v<2>, g, b, 0, s, src, pw, dst, lw, lh, scale_mode); break;
    case 4: hipLaunchKernelGGL(k_downsample_v<4>, g, b, 0, s, src, pw, dst, lw, lh, scale_mode); break;
    default: hipLaunchKernelGGL(k_downsample, g, b, 0, s, src, pw, dst, lw, lh, ch, scale_mode); break;
    }
    return hipGetLastError();
}

hipError_t pt_launch_downsample_vote(const int *src, int pw, int *dst, int lw, int lh, hipStream_t s)
{
    PT_SINK_REFUSE("object-ID pyramid");
    hipLaunchKernelGGL(k_downsample_vote, grid_rows(lw, lh), kRowBlock, 0, s, src, pw, dst, lw, lh);
    return hipGetLastError();
}

hipError_t pt_launch_fill(float *dst, size_t n, float v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (g_sink) {
        PtTask t = task_new(PT_TASK_FILL, (int64_t)((n + PT_TASK_ELEMS - 1) / PT_TASK_ELEMS), 0);
        t.q = dst;
        t.n[0] = (int64_t)n;
        t.f = v;
        return sink_push(t);
    }
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, n, v);
    return hipGetLastError();
}

hipError_t pt_launch_tensor_angle(const float *t, float *a, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (g_sink) {
        PtTask k = task_new(PT_TASK_TENSOR_ANGLE, (int64_t)((n + 255) / 256), 0);
        k.p[0] = t;
        k.q = a;
        k.n[0] = (int64_t)n;
        return sink_push(k);
    }
    hipLaunchKernelGGL(k_tensor_angle, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t, a, n);
    return hipGetLastError();
}

hipError_t pt_launch_interleave(const float4 *src, int w, int h, float4 *ex, int ew, int slot, hipStream_t s)
{
    if (g_sink) {
        PtTask t = task_2d(PT_TASK_INTERLEAVE, w, h, 16, 16);
        t.p[0] = src;
        t.q = ex;
        t.i[0] = w;
        t.i[1] = h;
        t.i[2] = ew;
        t.i[3] = slot;
        return sink_push(t);
    }
    dim3 b(256), g((w + 15) / 16, (h + 15) / 16);
    hipLaunchKernelGGL(k_interleave, g, b, 0, s, src, w, h, ex, ew, slot);
    return hipGetLastError();
}

hipError_t pt_launch_random_pyramid(const float4 *color, const float *iorient, int w, int h, const float4 *eout,
                                    const float *eorient, int ew, int eh, int transp, int orient_en,
                                    uint32_t key0, uint32_t key1, float4 *out, hipStream_t s)
{
    if (g_sink) {
        PtTask t = task_2d(PT_TASK_RANDOM_PYRAMID, w, h, 16, 16);
        t.p[0] = color;
        t.p[1] = iorient;
        t.p[2] = eout;
        t.p[3] = eorient;
        t.q = out;
        t.i[0] = w;
        t.i[1] = h;
        t.i[2] = ew;
        t.i[3] = eh;
        t.i[4] = transp;
        t.i[5] = orient_en;
        t.u[0] = key0;
        t.u[1] = key1;
        return sink_push(t);
    }
    dim3 b(256), g((w + 15) / 16, (h + 15) / 16);
    hipLaunchKernelGGL(k_random_pyramid, g, b, 0, s, color, iorient, w, h, eout, eorient, ew, eh, transp,
                       orient_en, key0, key1, out);
    return hipGetLastError();
}

hipError_t pt_launch_pack_nnf(const PtNnf &src, int W, int lw, int lh, void *dst, hipStream_t s)
{
    if (g_sink) {
        PtTask t = task_2d(PT_TASK_PACK_NNF, lw, lh, 16, 16);
        t.p[0] = src.xyth;
        t.p[1] = src.sc;
        t.p[2] = src.lum;
        t.p[3] = src.la;
        t.q = dst;
        t.i[0] = W;
        t.i[1] = lw;
        t.i[2] = lh;
        return sink_push(t);
    }
    dim3 b(256), g((lw + 15) / 16, (lh + 15) / 16);
    hipLaunchKernelGGL(k_pack_nnf, g, b, 0, s, src, W, lw, lh, (PatchXF36 *)dst);
    return hipGetLastError();
}

hipError_t pt_launch_unpack_nnf(const void *src, int lw, int lh, PtNnf dst, int W, hipStream_t s)
{
    PT_SINK_REFUSE("offset unpack");
    dim3 b(256), g((lw + 15) / 16, (lh + 15) / 16);
    hipLaunchKernelGGL(k_unpack_nnf, g, b, 0, s, (const PatchXF36 *)src, lw, lh, dst, W);
    return hipGetLastError();
}

/* ------------------------------------------------------------------ */
/* copies of recorded stream work                                      */
/* ------------------------------------------------------------------ */
hipError_t pt_copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                     hipMemcpyKind kind, hipStream_t s)
{
    if (!g_sink) return hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, kind, s);
    if (width == 0 || height == 0) return hipSuccess;
    if (kind != hipMemcpyDeviceToDevice) return sink_refuse("host copy");
    /* the widest element (16, 8 or 4 bytes) that divides every address,
     * pitch and row width */
    uint64_t m = (uint64_t)(uintptr_t)dst | (uint64_t)(uintptr_t)src | (uint64_t)width |
                 (height > 1 ? ((uint64_t)dpitch | (uint64_t)spitch) : 0);
    int e = (m % 16 == 0) ? 16 : ((m % 8 == 0) ? 8 : ((m % 4 == 0) ? 4 : 0));
    if (e == 0) return sink_refuse("unaligned copy");
    int64_t per_row = (int64_t)(width / e);
    int64_t tx = (per_row + PT_TASK_ELEMS - 1) / PT_TASK_ELEMS;
    if (tx * (int64_t)height > (1ll << 30) || per_row > (1ll << 30)) return sink_refuse("oversized copy");
    PtTask t = task_new(PT_TASK_COPY, tx * (int64_t)height, (int)tx);
    t.variant = e;
    t.q = dst;
    t.p[0] = src;
    t.n[0] = (int64_t)dpitch;
    t.n[1] = (int64_t)spitch;
    t.i[0] = (int)per_row;
    t.i[1] = (int)height;
    return sink_push(t);
}

hipError_t pt_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s)
{
    if (!g_sink) return hipMemcpyAsync(dst, src, bytes, kind, s);
    return pt_copy2d(dst, bytes, src, bytes, bytes, 1, kind, s);
}

/* ------------------------------------------------------------------ */
/* persistent batch kernel                                             */
/* ------------------------------------------------------------------ */
/*
 * One launch runs every step of every job of a batch.  Each job's program
 * is its plain path's sequence of kernels (PtTask: kind, operands, tiles);
 * the steps of one job run in order, each one only after every tile of the
 * previous one has completed, while steps of different jobs overlap freely.
 *
 * A job runs on a fixed set of XCDs; each step's tiles are split into one
 * contiguous band per XCD (row bands of the image, the same for every step),
 * so most of a step's inputs were written by the same XCD in the previous
 * step.  Within an XCD a tile's writes are complete for the next step once
 * they have reached its L2 (the vector L1 writes through): a claim completes
 * after a workgroup-scope release and s_waitcnt 0, and starts with an
 * agent-scope acquire (L1 invalidate).  When a job spans several XCDs, the
 * workgroup that completes an XCD's band of a step also writes that XCD's
 * L2 back (agent-scope release) before counting the band done, once per band
 * and step: a write-back at every claim measured slower than the work itself
 * (the batch took 1.7x as long).  The host places the jobs by their work
 * (longest first onto the least loaded XCDs; a job with more than an XCD's
 * share of the batch gets several), estimated from the patch taps of their
 * steps or measured in an earlier launch (P.cost); workgroups read the XCD
 * they run on from the hardware.
 *
 * Per job and band one 64-bit claim word holds (step, claims of the band in
 * the step, claims taken); a claim is a run of PtTask::chunk tiles.  A
 * workgroup's first wave reads the words of 64 jobs for its own XCD at once
 * and takes the first job (lowest index = highest priority, the host orders
 * jobs by work) with an open claim, with one atomic add; the add returns the
 * step and claim count it claimed in, so a claim that races with the opening
 * of the next step is a valid claim of that step.  The workgroup runs the
 * tiles with the same tile bodies as the kernels of the plain path, then adds
 * their number to the band's completed-tile count; the add that reaches the
 * band's cumulative end counts the band done, and the count that reaches the
 * step's cumulative number of bands writes the next step's claim words.
 *
 * A workgroup never waits for a particular tile: when no job has claimable
 * tiles it sleeps and looks again, and it exits when every job has finished.
 * A workgroup that sees no step complete anywhere for timeout_ticks of the
 * constant wall clock sets the give-up flag; every workgroup then exits, and
 * the host reports the failure.
 */
typedef const __attribute__((address_space(4))) PtTask ConstTask;
typedef const __attribute__((address_space(4))) int ConstInt;

template <class T> constexpr size_t pt_lds_max() { return sizeof(T); }
template <class T, class U, class... R> constexpr size_t pt_lds_max()
{
    return sizeof(T) > pt_lds_max<U, R...>() ? sizeof(T) : pt_lds_max<U, R...>();
}
/* the arena holds the LDS of every tile body the kernel runs */
constexpr size_t kPersistLds =
    pt_lds_max<PropUpLds<2, 8, 8, false, true>, PropUpLds<3, 8, 8, false, true>, PropUpLds<2, 8, 8, false, false>,
               PropUpLds<3, 8, 8, false, false>, ScatUpLds<2, 16, 8, false>, ScatUpLds<3, 16, 8, false>,
               ScatUpLds<2, 4, 4, false>, ScatUpLds<3, 4, 4, false>, PropUpLds<2, 16, 8, false, false>,
               PropUpLds<3, 16, 8, false, false>, ScatUpLds<2, 8, 8, false>, ScatUpLds<3, 8, 8, false>,
               ResLds<2, false>, ResLds<3, false>,
               ResTapLds<2, false>, ResTapLds<3, false>, PropTapLds<2, false>, PropTapLds<3, false>>();

/* Each tile body of the NNF operators runs as a function of its own, not
 * inlined into the claim loop: inlined, the loop keeps the lane-index
 * arithmetic and constants of every body live across all of them (measured
 * 151-256 VGPRs against 101-113 for the same bodies in their own kernels).
 * Called, each body has its own register allocation under the kernel's
 * occupancy target (4 waves per SIMD, what the LDS arena allows anyway), at
 * the cost of saving callee-saved registers once per tile.  Arguments are
 * made wave-uniform inside, so argument-table loads stay scalar and the
 * arena stays in LDS. */
typedef __attribute__((address_space(3))) char *LdsArena;

__device__ __forceinline__ ConstArgs *uniform_args(ConstArgs *p)
{
    uint64_t v = (uint64_t)(uintptr_t)p;
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (ConstArgs *)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ LdsArena uniform_arena(LdsArena p)
{
    return (LdsArena)(uintptr_t)__builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)p);
}

#define PT_PERSIST_BODY(NAME, TPL, CALL)                                                                  \
    TPL __device__ __attribute__((noinline)) void NAME(ConstArgs *a_, LdsArena ar_, int bx, int by, int tile) \
    {                                                                                                     \
        const PtArgs &A = (const PtArgs &)*uniform_args(a_);                                              \
        void *arena = (void *)uniform_arena(ar_);                                                         \
        bx = __builtin_amdgcn_readfirstlane(bx);                                                          \
        by = __builtin_amdgcn_readfirstlane(by);                                                          \
        tile = __builtin_amdgcn_readfirstlane(tile);                                                      \
        CALL;                                                                                             \
    }

#define PT_COMMA ,
PT_PERSIST_BODY(pk_prop_up, template <int RT PT_COMMA bool EE PT_COMMA bool XL>,
                (b_propagate_up<RT, EE, 8, 8, false, XL, true>(A, bx, by, A.jump, arena)))
PT_PERSIST_BODY(pk_prop_ee_a, template <int RT>,
                (b_propagate_up<RT, true, 16, 8, false, false, true>(A, bx, by, A.jump, arena)))
PT_PERSIST_BODY(pk_scat_ee_b, template <int RT>, (b_scatter_up<RT, true, 8, 8, false, true>(A, bx, by, A.nscat, arena)))
PT_PERSIST_BODY(pk_prop_tp, template <int RT>, (b_propagate_tp<RT, false, true>(A, tile, A.jump, arena)))
PT_PERSIST_BODY(pk_scat_ee, template <int RT>, (b_scatter_up<RT, true, 16, 8, false, true>(A, bx, by, A.nscat, arena)))
PT_PERSIST_BODY(pk_scat_a, template <int RT>, (b_scatter_up<RT, false, 4, 4, false, true>(A, bx, by, A.nscat, arena)))
PT_PERSIST_BODY(pk_res, template <int RT>, (b_update_residual<RT, false, false, true>(A, bx, by, arena)))
PT_PERSIST_BODY(pk_res_tp, template <int RT>, (b_update_residual_tp<RT, false, true>(A, tile, arena)))
#undef PT_COMMA
#undef PT_PERSIST_BODY

template <int RT>
__device__ __forceinline__ void persist_nnf(const PtTask &t, ConstArgs *A, int bx, int by, int tile, LdsArena arena)
{
    const int shape = t.variant >> 4;
    const bool xl = (t.variant & 8) != 0;
    if (t.kind == PT_TASK_PROPAGATE) {
        if (shape == PT_SHAPE_TAP) pk_prop_tp<RT>(A, arena, bx, by, tile);
        else if (shape == PT_SHAPE_UP_EE_A) pk_prop_ee_a<RT>(A, arena, bx, by, tile);
        else if (shape == PT_SHAPE_UP_EE_B) {
            if (xl) pk_prop_up<RT, true, true>(A, arena, bx, by, tile);
            else pk_prop_up<RT, true, false>(A, arena, bx, by, tile);
        } else {
            if (xl) pk_prop_up<RT, false, true>(A, arena, bx, by, tile);
            else pk_prop_up<RT, false, false>(A, arena, bx, by, tile);
        }
    } else if (t.kind == PT_TASK_SCATTER) {
        if (shape == PT_SHAPE_UP_EE_A) pk_scat_ee<RT>(A, arena, bx, by, tile);
        else if (shape == PT_SHAPE_UP_EE_B) pk_scat_ee_b<RT>(A, arena, bx, by, tile);
        else pk_scat_a<RT>(A, arena, bx, by, tile);
    } else {
        if (shape == PT_SHAPE_TAP) pk_res_tp<RT>(A, arena, bx, by, tile);
        else pk_res<RT>(A, arena, bx, by, tile);
    }
}

template <class V>
__device__ __forceinline__ void persist_copy(const PtTask &t, int row, int seg)
{
    const V *src = (const V *)((const char *)t.p[0] + (int64_t)row * t.n[1]);
    V *dst = (V *)((char *)t.q + (int64_t)row * t.n[0]);
    const int e0 = seg * PT_TASK_ELEMS + threadIdx.x;
#pragma unroll
    for (int k = 0; k < PT_TASK_ELEMS / 256; k++) {
        int e = e0 + k * 256;
        if (e < t.i[0]) dst[e] = src[e];
    }
}

__device__ __forceinline__ void persist_tile(const PtPersist &P, const PtTask &t, int tile, void *arena)
{
    const int bx = tile % t.tiles_x, by = tile / t.tiles_x;
    switch (t.kind) {
    case PT_TASK_PROPAGATE:
    case PT_TASK_SCATTER:
    case PT_TASK_RESIDUAL: {
        ConstArgs *A = (ConstArgs *)P.args + t.arg;
        if ((t.variant & 7) == 2) persist_nnf<2>(t, A, bx, by, tile, (LdsArena)arena);
        else persist_nnf<3>(t, A, bx, by, tile, (LdsArena)arena);
        break;
    }
    case PT_TASK_SYNC: {
        const PtArgs &A = (const PtArgs &)((ConstArgs *)P.args)[t.arg];
        b_sync_canvas(A, bx, by, A.dw, A.dh);
        break;
    }
    case PT_TASK_COPY_RANDOM: b_copy_random((const PtArgs &)((ConstArgs *)P.args)[t.arg], bx, by); break;
    case PT_TASK_UPSAMPLE: b_upsample((const PtArgs &)((ConstArgs *)P.args)[t.arg], bx, by); break;
    case PT_TASK_COPY:
        if (t.variant == 16) persist_copy<float4>(t, by, bx);
        else if (t.variant == 8) persist_copy<float2>(t, by, bx);
        else persist_copy<float>(t, by, bx);
        break;
    case PT_TASK_FILL: {
        float *dst = (float *)t.q;
#pragma unroll
        for (int k = 0; k < PT_TASK_ELEMS / 256; k++) {
            int64_t e = (int64_t)tile * PT_TASK_ELEMS + k * 256 + threadIdx.x;
            if (e < t.n[0]) dst[e] = t.f;
        }
        break;
    }
    case PT_TASK_DOWNSAMPLE: {
        const float *src = (const float *)t.p[0];
        float *dst = (float *)t.q;
        if (t.variant == 1) b_downsample_v<1>(src, t.i[0], dst, t.i[1], t.i[2], t.i[3], bx, by);
        else if (t.variant == 2) b_downsample_v<2>(src, t.i[0], dst, t.i[1], t.i[2], t.i[3], bx, by);
        else if (t.variant == 4) b_downsample_v<4>(src, t.i[0], dst, t.i[1], t.i[2], t.i[3], bx, by);
        else b_downsample(src, t.i[0], dst, t.i[1], t.i[2], t.variant, t.i[3], bx, by);
        break;
    }
    case PT_TASK_TENSOR_ANGLE: {
        int64_t e = (int64_t)tile * 256 + threadIdx.x;
        const float *tn = (const float *)t.p[0];
        if (e < t.n[0]) ((float *)t.q)[e] = pt_tensor_to_angle(tn[3 * e], tn[3 * e + 1], tn[3 * e + 2]);
        break;
    }
    case PT_TASK_INTERLEAVE:
        b_interleave((const float4 *)t.p[0], t.i[0], t.i[1], (float4 *)t.q, t.i[2], t.i[3], bx, by);
        break;
    case PT_TASK_RANDOM_PYRAMID:
        b_random_pyramid((const float4 *)t.p[0], (const float *)t.p[1], t.i[0], t.i[1], (const float4 *)t.p[2],
                         (const float *)t.p[3], t.i[2], t.i[3], t.i[4], t.i[5], t.u[0], t.u[1], (float4 *)t.q, bx,
                         by);
        break;
    case PT_TASK_PACK_NNF: {
        PtNnf n;
        n.xyth = (float4 *)t.p[0];
        n.sc = (float2 *)t.p[1];
        n.lum = (float *)t.p[2];
        n.la = (int32_t *)t.p[3];
        b_pack_nnf(n, t.i[0], t.i[1], t.i[2], (PatchXF36 *)t.q, bx, by);
        break;
    }
    default: break;
    }
}

/* 4 waves per SIMD: four workgroups of 256 lanes per CU, as many as the LDS
 * arena allows */
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_persistent(PtPersist P)
{
    __shared__ __attribute__((aligned(16))) char arena[kPersistLds];
    __shared__ int s_mode, s_job, s_step, s_claim, s_band;
    ConstTask *tasks = (ConstTask *)P.tasks;
    ConstInt *first = (ConstInt *)P.first;
    /* the XCD this workgroup runs on (HW_REG_XCC_ID, bits 3:0) */
    const int xcd = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);
    int seen = -1; /* wave 0, lane 0: completed steps at the last look */
    long long quiet_since = 0;
    for (;;) {
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            int job = -1, step = 0, claim = 0, band = 0;
            /* the give-up flag (set by a workgroup that timed out, or by the
             * host before the launch) stops every workgroup at its next look */
            const int gave =
                __builtin_amdgcn_readfirstlane(__hip_atomic_load(P.ctl + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            for (int base = 0; base < P.njobs && job < 0 && !gave; base += 64) {
                const int j = base + lane;
                unsigned long long w = 0;
                int b = 0;
                bool ok = false;
                if (j < P.njobs) {
                    const int xm = first[P.njobs + 1 + j];
                    if ((xm >> xcd) & 1) {
                        b = __builtin_popcount(xm & ((1 << xcd) - 1)); /* this XCD's band */
                        w = __hip_atomic_load(P.claim + (size_t)j * 8 + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = pt_cw_taken(w) < pt_cw_claims(w);
                    }
                }
                unsigned long long m = __ballot(ok);
                while (m != 0ull && job < 0) {
                    const int l = __builtin_ctzll(m);
                    m &= m - 1ull;
                    int gs = -1, gk = 0;
                    if (lane == l) {
                        const unsigned long long old = __hip_atomic_fetch_add(
                            P.claim + (size_t)j * 8 + b, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (pt_cw_taken(old) < pt_cw_claims(old)) {
                            gs = pt_cw_step(old);
                            gk = pt_cw_taken(old);
                        }
                    }
                    gs = __shfl(gs, l);
                    if (gs >= 0) {
                        job = base + l;
                        step = gs;
                        claim = __shfl(gk, l);
                        band = __shfl(b, l);
                    }
                }
            }
            if (job >= 0) {
                /* this CU's L1 may hold lines the job's earlier steps rewrote */
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            if (lane == 0) {
                int mode = 0;
                if (job < 0) {
                    const int fin = __hip_atomic_load(P.ctl + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    mode = (fin >= P.njobs || gave) ? 2 : 1;
                    if (mode == 1) {
                        const int prog = __hip_atomic_load(P.ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        const long long now = wall_clock64();
                        if (prog != seen) {
                            seen = prog;
                            quiet_since = now;
                        } else if (now - quiet_since > P.timeout_ticks) {
                            __hip_atomic_store(P.ctl + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            mode = 2;
                        }
                    }
                }
                s_mode = mode;
                s_job = job;
                s_step = step;
                s_claim = claim;
                s_band = band;
            }
        }
        __syncthreads();
        const int mode = s_mode, job = s_job, step = s_step, claim = s_claim, band = s_band;
        __syncthreads();
        if (mode == 2) break;
        if (mode == 1) {
            __builtin_amdgcn_s_sleep(32);
            continue;
        }
        const PtTask &t = (const PtTask &)tasks[first[job] + step];
        const int t0 = t.band[band] + claim * t.chunk, t1 = min(t0 + t.chunk, t.band[band + 1]);
        const long long c0 = clock64();
        for (int k = t0; k < t1; k++) {
            persist_tile(P, t, k, arena);
            __syncthreads();
        }
        const long long c1 = clock64();
        /* every wave's writes have reached the XCD's L2 before the count below */
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int cnt = t1 - t0;
            __hip_atomic_fetch_add(P.cost + job, (unsigned long long)(c1 - c0), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            const int old =
                __hip_atomic_fetch_add(P.done + (size_t)job * 8 + band, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old + cnt == t.band_end[band]) {
                /* this XCD's band of the step is complete */
                const int xm = first[P.njobs + 1 + job];
                const int nb = __builtin_popcount(xm);
                if (nb > 1) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); /* write this XCD's L2 back */
                const int parts = __hip_atomic_fetch_add(P.parts + job, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (parts + 1 == t.parts_end) {
                    /* the step is complete: open the next one on every band */
                    if (nb > 1) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                    const int ns = first[job + 1] - first[job];
                    for (int bb = 0; bb < nb; bb++) {
                        unsigned long long w = pt_claim_word(ns, 0);
                        if (step + 1 < ns) {
                            const PtTask &u = (const PtTask &)tasks[first[job] + step + 1];
                            w = pt_claim_word(step + 1, (u.band[bb + 1] - u.band[bb] + u.chunk - 1) / u.chunk);
                        }
                        __hip_atomic_store(P.claim + (size_t)job * 8 + bb, w, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    }
                    if (step + 1 == ns) __hip_atomic_fetch_add(P.ctl + 0, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_fetch_add(P.ctl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
}

int pt_task_op(int op, int shape, const PtArgs &a, int arg, PtTask *t, const char **why)
{
    *why = nullptr;
    if (op < 0 || op >= PT_OP_COUNT) {
        *why = "unknown operator";
        return PT_ERR_UNSUPPORTED;
    }
    if (op != PT_OP_SYNC_CANVAS) {
        if (a.time_active) *why = "temporal terms";
        else if (ext_terms(a)) *why = "extended patch-distance terms or style layers";
        else if (a.keys) *why = "keyframe search";
        else if (a.r != 2 && a.r != 3) *why = "patch radius other than 2 or 3";
        if (*why) return PT_ERR_UNSUPPORTED;
    }
    if (!pt_shape_valid(op, shape, a)) {
        *why = "invalid kernel shape";
        return PT_ERR_UNSUPPORTED;
    }
    int v = 0;
    switch (op) {
    case PT_OP_PROPAGATE:
        if (shape != PT_SHAPE_TAP && shape != PT_SHAPE_UP_EE_A && shape != PT_SHAPE_UP_EE_B && shape != PT_SHAPE_UP_A)
            *why = "propagate shape";
        v = (shape << 4) | (prop_xl(shape, a) ? 8 : 0) | a.r;
        break;
    case PT_OP_SCATTER:
        if (shape != PT_SHAPE_UP_EE_A && shape != PT_SHAPE_UP_EE_B && shape != PT_SHAPE_UP_A) *why = "scatter shape";
        v = (shape << 4) | a.r;
        break;
    case PT_OP_RESIDUAL:
        /* the refresh runs without the exemplar window (identical results) */
        if (shape != PT_SHAPE_TAP && shape != PT_SHAPE_PIXEL) *why = "refresh shape";
        v = (shape << 4) | a.r;
        break;
    default: break;
    }
    if (*why) return PT_ERR_UNSUPPORTED;
    int tx = 1;
    int n = pt_shape_tiles(op, shape, a, &tx);
    *t = task_new(op, n, tx);
    t->variant = v;
    t->arg = arg;
    return PT_OK;
}

/* Each workgroup records the XCD it runs on. */
__global__ void k_xcd_probe(unsigned *mask)
{
    if (threadIdx.x == 0) atomicOr(mask, 1u << (__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15));
}

hipError_t pt_launch_xcd_probe(int grid, unsigned *mask, hipStream_t s)
{
    if (grid < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_xcd_probe, dim3(grid), dim3(64), 0, s, mask);
    return hipGetLastError();
}

int pt_persistent_grid(int device)
{
    int nb = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_persistent, 256, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    return nb * cus;
}

hipError_t pt_launch_persistent(const PtPersist &p, int grid, hipStream_t s)
{
    if (grid < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_persistent, dim3(grid), dim3(256), 0, s, p);
    return hipGetLastError();
}
