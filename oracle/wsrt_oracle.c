/*
 * wsrt_oracle.c - CPU oracle of wgpu-spectral-rt's render loop.
 *
 * TEST INFRASTRUCTURE ONLY (see wsrt_oracle.h). Each function cites the
 * reference file:line it restates. f32 throughout, compiled with
 * -ffp-contract=off so that the order of operations is the WGSL order.
 */
#include "wsrt_oracle.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>

#include "or_math.h"

/* ====================================================================== */
/* KAT entry points                                                        */
/* ====================================================================== */
uint32_t or_pcg_hash(uint32_t x) { return or_pcg(x); }

void or_rand_stream(uint32_t seed_input, uint32_t n, float* out) {
    uint32_t s = or_pcg(seed_input); /* init_rng: shader.wgsl:146-149 */
    for (uint32_t i = 0; i < n; ++i) out[i] = or_rand(&s);
}
void or_cie_xyz_f(float l, float* o) {
    v3 c = or_cie_xyz(l);
    o[0] = c.x; o[1] = c.y; o[2] = c.z;
}
float or_rgb_to_spectrum_f(const float* rgb, float l) {
    return or_rgb_to_spectrum(V3(rgb[0], rgb[1], rgb[2]), l);
}
float or_blackbody_f(float l, float t) { return or_blackbody(l, t); }
float or_thin_film_f(float c, float n1, float n2, float n3, float d, float l) {
    return or_thin_film(c, n1, n2, n3, d, l);
}
float or_schlick_f(float c, float r) { return or_reflectance(c, r); }
void or_cauchy_constants(float* cmin, float* crange, float* pnorm) {
    or_cauchy_cdf(cmin, crange);
    *pnorm = 1.0f / (*crange * OR_PI * 40.0f);
}
float or_cauchy_sample_f(float u) {
    float cmin, crange;
    or_cauchy_cdf(&cmin, &crange);
    return or_cauchy_sample(u, cmin, crange);
}
void or_grid_jitter_f(uint32_t frame, float cell, float* o) {
    v3 j = or_grid_jitter(frame, cell);
    o[0] = j.x; o[1] = j.y; o[2] = j.z;
}
uint32_t or_hash_position_f(const float* p, uint32_t frame, float cell, uint32_t hs) {
    return or_hash_cell(V3(p[0], p[1], p[2]), or_grid_jitter(frame, cell), cell, hs);
}

/* shader.wgsl:734-767 */
static v4 hero_wavelengths(float u_rand, float rot) {
    float cmin, crange;
    or_cauchy_cdf(&cmin, &crange);
    float hero = or_cauchy_sample(u_rand, cmin, crange);
    const float q = OR_LAMBDA_RANGE / 4.0f;
    v4 w = V4(hero, hero + q * 1.0f, hero + q * 2.0f, hero + q * 3.0f);
    w.x = OR_LAMBDA_MIN + or_fmod_wgsl(w.x - OR_LAMBDA_MIN, OR_LAMBDA_RANGE);
    w.y = OR_LAMBDA_MIN + or_fmod_wgsl(w.y - OR_LAMBDA_MIN, OR_LAMBDA_RANGE);
    w.z = OR_LAMBDA_MIN + or_fmod_wgsl(w.z - OR_LAMBDA_MIN, OR_LAMBDA_RANGE);
    w.w = OR_LAMBDA_MIN + or_fmod_wgsl(w.w - OR_LAMBDA_MIN, OR_LAMBDA_RANGE);
    if (rot < 0.25f) {
    } else if (rot < 0.5f) {
        w = V4(w.y, w.z, w.w, w.x);
    } else if (rot < 0.75f) {
        w = V4(w.z, w.w, w.x, w.y);
    } else {
        w = V4(w.w, w.x, w.y, w.z);
    }
    return w;
}
void or_hero_wavelengths(float u, float rot, float* o) {
    v4 w = hero_wavelengths(u, rot);
    o[0] = w.x; o[1] = w.y; o[2] = w.z; o[3] = w.w;
}

/* ====================================================================== */
/* Geometry: src/scene/geometry.rs                                         */
/* ====================================================================== */
static void set_vertex(wsrt_vertex* v, float px, float py, float pz, float nx, float ny, float nz) {
    v->pos[0] = px; v->pos[1] = py; v->pos[2] = pz; v->pos[3] = 1.0f;
    v->normal[0] = nx; v->normal[1] = ny; v->normal[2] = nz; v->normal[3] = 0.0f;
}

/* geometry.rs:12-35 */
int or_geometry_plane(wsrt_vertex* v, uint32_t* idx, uint32_t* nv, uint32_t* ni) {
    *nv = 4; *ni = 6;
    if (!v || !idx) return 0;
    set_vertex(&v[0], -0.5f, 0.0f, 0.5f, 0, 1, 0);
    set_vertex(&v[1], 0.5f, 0.0f, 0.5f, 0, 1, 0);
    set_vertex(&v[2], -0.5f, 0.0f, -0.5f, 0, 1, 0);
    set_vertex(&v[3], 0.5f, 0.0f, -0.5f, 0, 1, 0);
    const uint32_t i[6] = {0, 1, 2, 2, 1, 3};
    memcpy(idx, i, sizeof i);
    return 0;
}

/* geometry.rs:38-118 */
int or_geometry_cube(wsrt_vertex* v, uint32_t* idx, uint32_t* nv, uint32_t* ni) {
    *nv = 24; *ni = 36;
    if (!v || !idx) return 0;
    static const float sides[6][5][3] = {
        {{0, 0, 1}, {-0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, 0.5f}, {0.5f, 0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f}},
        {{0, 0, -1}, {0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, -0.5f}, {-0.5f, 0.5f, -0.5f}, {0.5f, 0.5f, -0.5f}},
        {{0, 1, 0}, {-0.5f, 0.5f, 0.5f}, {0.5f, 0.5f, 0.5f}, {0.5f, 0.5f, -0.5f}, {-0.5f, 0.5f, -0.5f}},
        {{0, -1, 0}, {-0.5f, -0.5f, -0.5f}, {0.5f, -0.5f, -0.5f}, {0.5f, -0.5f, 0.5f}, {-0.5f, -0.5f, 0.5f}},
        {{1, 0, 0}, {0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, -0.5f}, {0.5f, 0.5f, -0.5f}, {0.5f, 0.5f, 0.5f}},
        {{-1, 0, 0}, {-0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f}, {-0.5f, 0.5f, -0.5f}},
    };
    uint32_t vi = 0, k = 0;
    for (int f = 0; f < 6; ++f) {
        const float* n = sides[f][0];
        for (int c = 1; c <= 4; ++c)
            set_vertex(&v[vi + c - 1], sides[f][c][0], sides[f][c][1], sides[f][c][2], n[0], n[1], n[2]);
        idx[k++] = vi; idx[k++] = vi + 1; idx[k++] = vi + 2;
        idx[k++] = vi; idx[k++] = vi + 2; idx[k++] = vi + 3;
        vi += 4;
    }
    return 0;
}

/* Icosphere (geometry.rs:121-232): 12 + 30*(4^s - 1)/3 ... vertices. */
typedef struct { uint64_t key; uint32_t val; } mp_slot;
typedef struct { mp_slot* slots; uint32_t cap; } mp_map;
static uint32_t* mp_get(mp_map* m, uint64_t key, int* found) {
    uint64_t h = key * 0x9E3779B97F4A7C15ull;
    for (uint32_t i = (uint32_t)(h >> 40) & (m->cap - 1);; i = (i + 1) & (m->cap - 1)) {
        if (m->slots[i].key == key + 1) { *found = 1; return &m->slots[i].val; }
        if (m->slots[i].key == 0) { m->slots[i].key = key + 1; *found = 0; return &m->slots[i].val; }
    }
}

static void sphere_push(wsrt_vertex* v, uint32_t* n, float px, float py, float pz) {
    float length = sqrtf(px * px + py * py + pz * pz);
    float nx = px / length, ny = py / length, nz = pz / length;
    set_vertex(&v[*n], nx * 0.5f, ny * 0.5f, nz * 0.5f, nx, ny, nz);
    (*n)++;
}

static uint32_t sphere_counts_v(uint32_t s) {
    uint32_t v = 12, e = 30, f = 20;
    for (uint32_t i = 0; i < s; ++i) { v += e; e = 2 * e + 3 * f; f *= 4; }
    return v;
}

int or_geometry_sphere(uint32_t subdiv, wsrt_vertex* v, uint32_t* idx, uint32_t* nv, uint32_t* ni) {
    if (subdiv > 8) return -1;
    uint32_t nf = 20u << (2 * subdiv);
    *nv = sphere_counts_v(subdiv);
    *ni = nf * 3;
    if (!v || !idx) return 0;
    float t = (1.0f + sqrtf(5.0f)) / 2.0f;
    uint32_t n = 0;
    const float base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                               {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (int i = 0; i < 12; ++i) sphere_push(v, &n, base[i][0], base[i][1], base[i][2]);
    static const uint32_t f0[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                                       {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                       {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                                       {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    uint32_t* faces = (uint32_t*)malloc(sizeof(uint32_t) * 3 * nf);
    uint32_t* nfaces = (uint32_t*)malloc(sizeof(uint32_t) * 3 * nf);
    memcpy(faces, f0, sizeof f0);
    uint32_t cur = 20;
    mp_map map;
    map.cap = 1;
    while (map.cap < 4 * (*nv)) map.cap <<= 1;
    map.slots = (mp_slot*)calloc(map.cap, sizeof(mp_slot));
    for (uint32_t s = 0; s < subdiv; ++s) {
        uint32_t k = 0;
        for (uint32_t f = 0; f < cur; ++f) {
            uint32_t tri[3] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
            uint32_t mid[3];
            for (int e = 0; e < 3; ++e) { /* (v1,v2), (v2,v3), (v3,v1) */
                uint32_t p1 = tri[e], p2 = tri[(e + 1) % 3];
                uint64_t key = p1 < p2 ? ((uint64_t)p1 << 32 | p2) : ((uint64_t)p2 << 32 | p1);
                int found;
                uint32_t* slot = mp_get(&map, key, &found);
                if (!found) { /* geometry.rs:201-232 */
                    const or_f32* a = v[p1].pos;
                    const or_f32* b = v[p2].pos;
                    float mx = (a[0] + b[0]) * 0.5f, my = (a[1] + b[1]) * 0.5f, mz = (a[2] + b[2]) * 0.5f;
                    sphere_push(v, &n, mx, my, mz);
                    *slot = n - 1;
                }
                mid[e] = *slot;
            }
            uint32_t a = mid[0], b = mid[1], c = mid[2];
            uint32_t nt[4][3] = {{tri[0], a, c}, {tri[1], b, a}, {tri[2], c, b}, {a, b, c}};
            for (int q = 0; q < 4; ++q)
                for (int r = 0; r < 3; ++r) nfaces[k++] = nt[q][r];
        }
        cur *= 4;
        uint32_t* tmp = faces; faces = nfaces; nfaces = tmp;
    }
    memcpy(idx, faces, sizeof(uint32_t) * 3 * cur);
    free(map.slots);
    free(faces);
    free(nfaces);
    return (n == *nv) ? 0 : -2;
}

/* ====================================================================== */
/* glam-style Mat4 (column-major), f32, glam's mul_vec4 summation order     */
/* ====================================================================== */
void or_mat4_identity(float* m) {
    memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}
static void mat4_mul_vec4(const float* a, const float* v, float* out) {
    float r[4];
    for (int i = 0; i < 4; ++i) {
        float s = a[0 + i] * v[0];
        s = s + a[4 + i] * v[1];
        s = s + a[8 + i] * v[2];
        s = s + a[12 + i] * v[3];
        r[i] = s;
    }
    memcpy(out, r, sizeof r);
}
void or_mat4_mul(const float* a, const float* b, float* out) {
    float r[16];
    for (int c = 0; c < 4; ++c) mat4_mul_vec4(a, b + 4 * c, r + 4 * c);
    memcpy(out, r, sizeof r);
}
void or_mat4_translation(float x, float y, float z, float* m) {
    or_mat4_identity(m);
    m[12] = x; m[13] = y; m[14] = z;
}
void or_mat4_scale(float x, float y, float z, float* m) {
    or_mat4_identity(m);
    m[0] = x; m[5] = y; m[10] = z;
}
void or_mat4_rotation_x(float a, float* m) {
    float s = sinf(a), c = cosf(a);
    or_mat4_identity(m);
    m[5] = c; m[6] = s; m[9] = -s; m[10] = c;
}
void or_mat4_rotation_y(float a, float* m) {
    float s = sinf(a), c = cosf(a);
    or_mat4_identity(m);
    m[0] = c; m[2] = -s; m[8] = s; m[10] = c;
}
void or_mat4_rotation_z(float a, float* m) {
    float s = sinf(a), c = cosf(a);
    or_mat4_identity(m);
    m[0] = c; m[1] = s; m[4] = -s; m[5] = c;
}

/* General 4x4 inverse by cofactor expansion (f64), rounded to f32. */
static void mat4_inverse(const float* mf, or_f32* out) {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = mf[i];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    for (int i = 0; i < 16; ++i) out[i] = (float)(inv[i] / det);
}

/* camera.rs:178-194: look_at_rh + perspective_rh(45 deg, aspect, 0.1, 100). */
void or_camera_uniform(const float* pos, float yaw, float pitch, float aspect, uint32_t frame,
                       wsrt_camera_uniform* out) {
    float sy = sinf(yaw), cy = cosf(yaw), sp = sinf(pitch), cp = cosf(pitch);
    v3 fwd = V3(cp * cy, sp, cp * sy);
    fwd = mul3s(fwd, 1.0f / sqrtf(dot3(fwd, fwd)));
    v3 eye = V3(pos[0], pos[1], pos[2]);
    v3 f = fwd; /* center - eye = fwd up to rounding; glam normalizes it again */
    v3 ctr = add3(eye, fwd);
    f = sub3(ctr, eye);
    f = mul3s(f, 1.0f / sqrtf(dot3(f, f)));
    v3 up = V3(0, 1, 0);
    v3 s = cross3(f, up);
    s = mul3s(s, 1.0f / sqrtf(dot3(s, s)));
    v3 u = cross3(s, f);
    float view[16] = {s.x, u.x, -f.x, 0, s.y, u.y, -f.y, 0, s.z, u.z, -f.z, 0,
                      -dot3(eye, s), -dot3(eye, u), dot3(eye, f), 1};
    float fov = 45.0f * (OR_PI / 180.0f);
    float sf = sinf(0.5f * fov), cf = cosf(0.5f * fov);
    float h = cf / sf, w = h / aspect, r = 100.0f / (0.1f - 100.0f);
    float proj[16] = {w, 0, 0, 0, 0, h, 0, 0, 0, 0, r, -1, 0, 0, r * 0.1f, 0};
    for (int i = 0; i < 16; ++i) {
        out->view[i] = view[i];
        out->proj[i] = proj[i];
    }
    mat4_inverse(view, out->view_inverse);
    mat4_inverse(proj, out->proj_inverse);
    out->frame_count = frame;
    out->pad[0] = out->pad[1] = out->pad[2] = 0;
}

/* ====================================================================== */
/* Scene: src/scene/builder.rs                                              */
/* ====================================================================== */
typedef struct {
    uint32_t mesh, mat;
    float m[16];
    float w2o[9]; /* inverse of the linear part, column-major */
} or_instance;

typedef struct {
    float bmin[3], bmax[3];
    uint32_t first; /* leaf: first index into bvh_order; inner: left child */
    uint32_t count; /* 0: inner node (right child = left + 1 not assumed) */
    uint32_t right;
} or_bnode;

struct or_scene {
    wsrt_vertex* verts; uint32_t nv, cv;
    uint32_t* idx; uint32_t ni, ci;
    wsrt_mesh_info* meshes; uint32_t* mesh_nv; uint32_t* mesh_ni; uint32_t nmesh, cmesh;
    wsrt_material* mats; uint32_t nmat, cmat;
    or_instance* inst; uint32_t ninst, cinst;
    wsrt_light_info* lights; uint32_t nlights, clights;
    /* built */
    uint32_t ntri;
    float* v0; float* e1; float* e2; /* ntri x 3 each */
    float* wv;                        /* world vertices, ntri x 9 */
    uint32_t* tri_inst; uint32_t* tri_prim;
    or_bnode* nodes; uint32_t nnodes;
    uint32_t* order;
    /* environment light (extension), equirect RGB, row 0 = +Y */
    float* env; uint32_t env_w, env_h;
};

#define GROW(ptr, n, cap, add)                                                   \
    do {                                                                         \
        if ((n) + (add) > (cap)) {                                               \
            uint32_t nc = (cap) ? (cap) : 16;                                     \
            while (nc < (n) + (add)) nc *= 2;                                     \
            ptr = realloc(ptr, sizeof(*(ptr)) * nc);                              \
            cap = nc;                                                            \
        }                                                                        \
    } while (0)

or_scene* or_scene_create(void) { return (or_scene*)calloc(1, sizeof(or_scene)); }
void or_scene_destroy(or_scene* s) {
    if (!s) return;
    free(s->verts); free(s->idx); free(s->meshes); free(s->mesh_nv); free(s->mesh_ni);
    free(s->mats); free(s->inst); free(s->lights); free(s->v0); free(s->e1); free(s->e2);
    free(s->tri_inst); free(s->tri_prim); free(s->nodes); free(s->order); free(s->wv);
    free(s->env);
    free(s);
}
uint32_t or_scene_add_material(or_scene* s, const wsrt_material* m) {
    GROW(s->mats, s->nmat, s->cmat, 1);
    s->mats[s->nmat] = *m;
    return s->nmat++;
}
/* builder.rs:76-91 */
uint32_t or_scene_add_mesh(or_scene* s, const wsrt_vertex* v, uint32_t nv, const uint32_t* idx,
                           uint32_t ni) {
    uint32_t cm = s->cmesh;
    GROW(s->meshes, s->nmesh, s->cmesh, 1);
    if (s->cmesh != cm) {
        s->mesh_nv = realloc(s->mesh_nv, sizeof(uint32_t) * s->cmesh);
        s->mesh_ni = realloc(s->mesh_ni, sizeof(uint32_t) * s->cmesh);
    }
    wsrt_mesh_info mi = {s->nv, s->ni, {0, 0}};
    s->meshes[s->nmesh] = mi;
    s->mesh_nv[s->nmesh] = nv;
    s->mesh_ni[s->nmesh] = ni;
    GROW(s->verts, s->nv, s->cv, nv);
    memcpy(s->verts + s->nv, v, sizeof(wsrt_vertex) * nv);
    s->nv += nv;
    GROW(s->idx, s->ni, s->ci, ni);
    memcpy(s->idx + s->ni, idx, sizeof(uint32_t) * ni);
    s->ni += ni;
    return s->nmesh++;
}
static void linear_inverse(const float* m, float* w) {
    double a = m[0], b = m[4], c = m[8], d = m[1], e = m[5], f = m[9], g = m[2], h = m[6], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double D = -(b * i - c * h), E = a * i - c * g, F = -(a * h - b * g);
    double G = b * f - c * e, H = -(a * f - c * d), I = a * e - b * d;
    double det = a * A + b * B + c * C;
    /* inverse rows: [A D G; B E H; C F I] / det ; stored column-major */
    double r[3][3] = {{A / det, D / det, G / det}, {B / det, E / det, H / det}, {C / det, F / det, I / det}};
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) w[col * 3 + row] = (float)r[row][col];
}
void or_scene_add_instance(or_scene* s, uint32_t mesh, uint32_t mat, const float* m16) {
    GROW(s->inst, s->ninst, s->cinst, 1);
    or_instance* in = &s->inst[s->ninst++];
    in->mesh = mesh; in->mat = mat;
    memcpy(in->m, m16, sizeof in->m);
    linear_inverse(m16, in->w2o);
}
/* builder.rs:101-129 */
void or_scene_add_light_instance(or_scene* s, uint32_t mesh, uint32_t mat, const float* m16,
                                 const float* params) {
    or_scene_add_instance(s, mesh, mat, m16);
    const wsrt_mesh_info* mi = &s->meshes[mesh];
    for (uint32_t i = 0; i + 2 < s->mesh_ni[mesh]; i += 3) {
        wsrt_light_info L;
        or_f32* dst[3] = {L.v0, L.v1, L.v2};
        for (int k = 0; k < 3; ++k) {
            const or_f32* p = s->verts[mi->vertex_offset + s->idx[mi->index_offset + i + k]].pos;
            float v[4] = {p[0], p[1], p[2], 1.0f}, r[4];
            mat4_mul_vec4(m16, v, r);
            dst[k][0] = r[0]; dst[k][1] = r[1]; dst[k][2] = r[2]; dst[k][3] = 1.0f;
        }
        for (int k = 0; k < 4; ++k) L.params[k] = params[k];
        GROW(s->lights, s->nlights, s->clights, 1);
        s->lights[s->nlights++] = L;
    }
}
void or_scene_add_point_light(or_scene* s, const float* pos, const float* params) {
    wsrt_light_info L;
    memset(&L, 0, sizeof L);
    L.v0[0] = pos[0]; L.v0[1] = pos[1]; L.v0[2] = pos[2]; L.v0[3] = 1.0f;
    for (int k = 0; k < 4; ++k) L.params[k] = params[k];
    L.params[0] = 1.0f;
    GROW(s->lights, s->nlights, s->clights, 1);
    s->lights[s->nlights++] = L;
}
uint32_t or_scene_num_triangles(const or_scene* s) { return s->ntri; }
uint32_t or_scene_num_lights(const or_scene* s) { return s->nlights; }
void or_scene_get_lights(const or_scene* s, wsrt_light_info* out) {
    memcpy(out, s->lights, sizeof(wsrt_light_info) * s->nlights);
}
void or_scene_get_world(const or_scene* s, float* tris, uint32_t* ids) {
    for (uint32_t t = 0; t < s->ntri; ++t) {
        if (tris) memcpy(tris + 9 * t, s->wv + 9 * t, 9 * sizeof(float));
        if (ids) {
            ids[3 * t] = s->tri_inst[t];
            ids[3 * t + 1] = s->inst[s->tri_inst[t]].mat;
            ids[3 * t + 2] = s->tri_prim[t];
        }
    }
}

/* --- oracle BVH (binary, median split), independent of the HIP BVH8 --- */
static float* g_cent; /* qsort context */
static int g_axis;
static int cmp_cent(const void* a, const void* b) {
    float ca = g_cent[3 * (*(const uint32_t*)a) + g_axis], cb = g_cent[3 * (*(const uint32_t*)b) + g_axis];
    if (ca < cb) return -1;
    if (ca > cb) return 1;
    uint32_t ia = *(const uint32_t*)a, ib = *(const uint32_t*)b;
    return ia < ib ? -1 : (ia > ib);
}
static uint32_t bvh_build_rec(or_scene* s, const float* wv, float* cent, uint32_t first, uint32_t count) {
    uint32_t id = s->nnodes++;
    or_bnode* n = &s->nodes[id];
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float cmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, cmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = first; i < first + count; ++i) {
        uint32_t t = s->order[i];
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) {
                float p = wv[9 * t + 3 * k + a];
                if (p < mn[a]) mn[a] = p;
                if (p > mx[a]) mx[a] = p;
            }
        for (int a = 0; a < 3; ++a) {
            float c = cent[3 * t + a];
            if (c < cmn[a]) cmn[a] = c;
            if (c > cmx[a]) cmx[a] = c;
        }
    }
    memcpy(n->bmin, mn, sizeof mn);
    memcpy(n->bmax, mx, sizeof mx);
    if (count <= 4) {
        n->first = first; n->count = count; n->right = 0;
        return id;
    }
    int axis = 0;
    float ext = cmx[0] - cmn[0];
    for (int a = 1; a < 3; ++a)
        if (cmx[a] - cmn[a] > ext) { ext = cmx[a] - cmn[a]; axis = a; }
    g_cent = cent; g_axis = axis;
    qsort(s->order + first, count, sizeof(uint32_t), cmp_cent);
    uint32_t half = count / 2;
    n->count = 0;
    uint32_t l = bvh_build_rec(s, wv, cent, first, half);
    uint32_t r = bvh_build_rec(s, wv, cent, first + half, count - half);
    s->nodes[id].first = l;
    s->nodes[id].right = r;
    return id;
}

/* World triangles (instance order, then mesh triangle order), transformed
 * with glam's mul_vec4 order (builder.rs:118-120 uses the same product for
 * lights). */
int or_scene_build(or_scene* s) {
    uint32_t ntri = 0;
    for (uint32_t i = 0; i < s->ninst; ++i) ntri += s->mesh_ni[s->inst[i].mesh] / 3;
    s->ntri = ntri;
    s->v0 = realloc(s->v0, sizeof(float) * 3 * (ntri + 1));
    s->e1 = realloc(s->e1, sizeof(float) * 3 * (ntri + 1));
    s->e2 = realloc(s->e2, sizeof(float) * 3 * (ntri + 1));
    s->tri_inst = realloc(s->tri_inst, sizeof(uint32_t) * (ntri + 1));
    s->tri_prim = realloc(s->tri_prim, sizeof(uint32_t) * (ntri + 1));
    free(s->wv);
    float* wv = (float*)malloc(sizeof(float) * 9 * (ntri + 1));
    s->wv = wv;
    float* cent = (float*)malloc(sizeof(float) * 3 * (ntri + 1));
    uint32_t t = 0;
    for (uint32_t i = 0; i < s->ninst; ++i) {
        const or_instance* in = &s->inst[i];
        const wsrt_mesh_info* mi = &s->meshes[in->mesh];
        uint32_t nt = s->mesh_ni[in->mesh] / 3;
        for (uint32_t p = 0; p < nt; ++p, ++t) {
            float w[3][4];
            for (int k = 0; k < 3; ++k) {
                const or_f32* pos = s->verts[mi->vertex_offset + s->idx[mi->index_offset + 3 * p + k]].pos;
                float v[4] = {pos[0], pos[1], pos[2], 1.0f};
                mat4_mul_vec4(in->m, v, w[k]);
            }
            for (int a = 0; a < 3; ++a) {
                s->v0[3 * t + a] = w[0][a];
                s->e1[3 * t + a] = w[1][a] - w[0][a];
                s->e2[3 * t + a] = w[2][a] - w[0][a];
                for (int k = 0; k < 3; ++k) wv[9 * t + 3 * k + a] = w[k][a];
                float lo = fminf(w[0][a], fminf(w[1][a], w[2][a]));
                float hi = fmaxf(w[0][a], fmaxf(w[1][a], w[2][a]));
                cent[3 * t + a] = (lo + hi) * 0.5f;
            }
            s->tri_inst[t] = i;
            s->tri_prim[t] = p;
        }
    }
    free(s->nodes);
    free(s->order);
    s->nodes = (or_bnode*)calloc(2 * ntri + 1, sizeof(or_bnode));
    s->order = (uint32_t*)malloc(sizeof(uint32_t) * (ntri + 1));
    for (uint32_t k = 0; k < ntri; ++k) s->order[k] = k;
    s->nnodes = 0;
    if (ntri) bvh_build_rec(s, wv, cent, 0, ntri);
    free(cent);
    return 0;
}

/* ====================================================================== */
/* Ray queries. Closest hit on (tmin, tmax) with ties broken by the lower   */
/* triangle id; triangle test = Moller-Trumbore, no culling; front face =   */
/* det > 0, i.e. dot(dir, cross(e1, e2)) < 0.                               */
/* ====================================================================== */
typedef struct { float t, u, v; uint32_t tri; int front; } hit_t;
#define NO_HIT 0xFFFFFFFFu

/* cross and dot with fused multiply-adds, each component rounded once per
 * fused step (the HIP path's triangle test uses the same forms and order) */
static inline v3 cross3_fma(v3 a, v3 b) {
    return V3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
static inline float dot3_fma(v3 a, v3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

static inline int tri_intersect(v3 o, v3 d, v3 v0, v3 e1, v3 e2, float* t, float* u, float* v,
                                int* front) {
    /* Moller-Trumbore; the barycentric tests on the unscaled numerators
     * with det's sign normalised, the division only for a triangle that
     * passes them (the HIP path's order of operations) */
    v3 pvec = cross3_fma(d, e2);
    float det = dot3_fma(e1, pvec);
    if (det == 0.0f) return 0;
    float sg = det > 0.0f ? 1.0f : -1.0f, adet = det * sg;
    v3 tvec = sub3(o, v0);
    float U = dot3_fma(tvec, pvec);
    if (U * sg < 0.0f || U * sg > adet) return 0;
    v3 qvec = cross3_fma(tvec, e1);
    float V = dot3_fma(d, qvec);
    if (V * sg < 0.0f || (U + V) * sg > adet) return 0;
    float inv = 1.0f / det;
    *t = dot3_fma(e2, qvec) * inv;
    *u = U * inv; *v = V * inv; *front = det > 0.0f;
    return 1;
}

static inline void tri_candidate(const or_scene* s, uint32_t t, v3 o, v3 d, float tmin, hit_t* best) {
    float tt, u, v;
    int fr;
    v3 v0 = V3(s->v0[3 * t], s->v0[3 * t + 1], s->v0[3 * t + 2]);
    v3 e1 = V3(s->e1[3 * t], s->e1[3 * t + 1], s->e1[3 * t + 2]);
    v3 e2 = V3(s->e2[3 * t], s->e2[3 * t + 1], s->e2[3 * t + 2]);
    if (!tri_intersect(o, d, v0, e1, e2, &tt, &u, &v, &fr)) return;
    if (!(tt > tmin)) return;
    if (tt < best->t || (tt == best->t && best->tri != NO_HIT && t < best->tri)) {
        best->t = tt; best->u = u; best->v = v; best->tri = t; best->front = fr;
    }
}

/* Direction components of magnitude < 1e-20 are replaced by +-1e-20 for the
 * box tests only (keeps slab values finite). */
static inline float safe_inv(float d) {
    if (fabsf(d) < 1e-20f) d = d < 0.0f ? -1e-20f : 1e-20f;
    return 1.0f / d;
}
#define BOX_PAD 1.00001f

static int box_hit(const float* mn, const float* mx, v3 o, v3 inv, float tmin, float tmax) {
    float t0x = (mn[0] - o.x) * inv.x, t1x = (mx[0] - o.x) * inv.x;
    float t0y = (mn[1] - o.y) * inv.y, t1y = (mx[1] - o.y) * inv.y;
    float t0z = (mn[2] - o.z) * inv.z, t1z = (mx[2] - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
    return tn <= tf * BOX_PAD;
}

static hit_t scene_closest(const or_scene* s, v3 o, v3 d, float tmin, float tmax, int brute) {
    hit_t best = {tmax, 0, 0, NO_HIT, 0};
    if (!s->ntri) return best;
    if (brute) {
        for (uint32_t t = 0; t < s->ntri; ++t) tri_candidate(s, t, o, d, tmin, &best);
        return best;
    }
    v3 inv = V3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const or_bnode* n = &s->nodes[stack[--sp]];
        if (!box_hit(n->bmin, n->bmax, o, inv, tmin, best.t)) continue;
        if (n->count) {
            for (uint32_t i = n->first; i < n->first + n->count; ++i)
                tri_candidate(s, s->order[i], o, d, tmin, &best);
        } else {
            stack[sp++] = n->right;
            stack[sp++] = n->first;
        }
    }
    return best;
}

static int scene_any(const or_scene* s, v3 o, v3 d, float tmin, float tmax, int brute) {
    if (!s->ntri) return 0;
    hit_t h = {tmax, 0, 0, NO_HIT, 0};
    if (brute) {
        for (uint32_t t = 0; t < s->ntri; ++t) {
            tri_candidate(s, t, o, d, tmin, &h);
            if (h.tri != NO_HIT) return 1;
        }
        return 0;
    }
    v3 inv = V3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const or_bnode* n = &s->nodes[stack[--sp]];
        if (!box_hit(n->bmin, n->bmax, o, inv, tmin, tmax)) continue;
        if (n->count) {
            for (uint32_t i = n->first; i < n->first + n->count; ++i) {
                tri_candidate(s, s->order[i], o, d, tmin, &h);
                if (h.tri != NO_HIT) return 1;
            }
        } else {
            stack[sp++] = n->right;
            stack[sp++] = n->first;
        }
    }
    return 0;
}

static inline uint32_t fbits(float f) { or_f32 g = (or_f32)f; uint32_t u; memcpy(&u, &g, 4); return u; }
static inline float bitsf(uint32_t u) { or_f32 f; memcpy(&f, &u, 4); return f; }

void or_closest_hit(const or_scene* s, const float* rays, uint32_t n, int brute, uint32_t* hits,
                    uint32_t* flags) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* r = rays + 8 * i;
        hit_t h = scene_closest(s, V3(r[0], r[1], r[2]), V3(r[4], r[5], r[6]), r[3], r[7], brute);
        int hit = h.tri != NO_HIT;
        hits[4 * i] = fbits(hit ? h.t : INFINITY);
        hits[4 * i + 1] = h.tri;
        hits[4 * i + 2] = fbits(hit ? h.u : 0.0f);
        hits[4 * i + 3] = fbits(hit ? h.v : 0.0f);
        flags[i] = hit ? (1u | (h.front ? 2u : 0u)) : 0u;
    }
}
void or_any_hit(const or_scene* s, const float* rays, uint32_t n, int brute, uint32_t* occ) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* r = rays + 8 * i;
        occ[i] = (uint32_t)scene_any(s, V3(r[0], r[1], r[2]), V3(r[4], r[5], r[6]), r[3], r[7], brute);
    }
}

/* --- BVH8 traversal over the HIP path's arrays (format in wsrt.h) ------- */

void or_bvh8_closest(const wsrt_bvh8_node* nodes, uint32_t n_nodes, const wsrt_bvh_tri* tris,
                     const float* rays, uint32_t n, uint32_t* hits, uint32_t* flags,
                     uint32_t* node_visits, uint32_t* tri_tests) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < (int64_t)n; ++r) {
        const float* ry = rays + 8 * r;
        v3 o = V3(ry[0], ry[1], ry[2]), d = V3(ry[4], ry[5], ry[6]);
        float tmin = ry[3];
        float best_t = ry[7];
        uint32_t best = NO_HIT;
        float bu = 0, bv = 0;
        int bfront = 0;
        uint32_t nv = 0, nt = 0;
        float inv[3] = {safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)};
        float oo[3] = {o.x, o.y, o.z};
        int oct = (d.x < 0.0f ? 1 : 0) | (d.y < 0.0f ? 2 : 0) | (d.z < 0.0f ? 4 : 0);
        uint32_t st_node[256];
        float st_t[256];
        int sp = 0;
        /* a ray with a NaN in its origin or direction meets no triangle (its
         * Moller-Trumbore values fail every test) and is not traversed */
        const float sum = o.x + o.y + o.z + d.x + d.y + d.z;
        if (n_nodes && sum == sum) { st_node[sp] = 0; st_t[sp] = tmin; sp++; }
        while (sp) {
            --sp;
            uint32_t ni = st_node[sp];
            if (st_t[sp] > best_t * BOX_PAD) continue;
            const wsrt_bvh8_node* nd = &nodes[ni];
            nv++;
            float as[3], an[3], fs[3], fo[3];
            for (int a = 0; a < 3; ++a) {
                as[a] = ldexpf(inv[a], nd->exp[a]); /* 2^exp * inv, exact */
                const float ao = (nd->origin[a] - oo[a]) * inv[a];
                /* bound of the f32 error of this axis' slab values, off its
                 * near planes and onto its far planes (padding folded into
                 * the far planes' constants): hit when near - e <= far *
                 * BOX_PAD + e (1 + BOX_PAD) on every axis and near - e <=
                 * best_t * BOX_PAD */
                const float e = (fabsf(as[a]) * 255.0f + fabsf(ao)) * 4.0e-7f;
                an[a] = ao - e;
                fs[a] = as[a] * BOX_PAD;
                fo[a] = fmaf(ao, BOX_PAD, e * (1.0f + BOX_PAD));
            }
            const float tcap = best_t * BOX_PAD;
            float tnear[8];
            int chit[8];
            for (int c = 0; c < 8; ++c) {
                float tn = tmin, tf = tcap;
                for (int a = 0; a < 3; ++a) {
                    int ng = inv[a] < 0.0f;
                    float qn = (float)(ng ? nd->qhi[a][c] : nd->qlo[a][c]);
                    float qf = (float)(ng ? nd->qlo[a][c] : nd->qhi[a][c]);
                    tn = fmaxf(tn, fmaf(qn, as[a], an[a]));
                    tf = fminf(tf, fmaf(qf, fs[a], fo[a]));
                }
                tnear[c] = tn;
                chit[c] = tn <= tf; /* an empty slot is a leaf of no triangles */
            }
            /* leaf children first, in slot order */
            for (int c = 0; c < 8; ++c) {
                if (!chit[c] || (nd->meta[c] & 0x80)) continue;
                uint32_t cnt = nd->meta[c] >> 5, off = nd->meta[c] & 0x1F;
                for (uint32_t k = 0; k < cnt; ++k) {
                    const wsrt_bvh_tri* tr = &tris[nd->tri_base + off + k];
                    nt++;
                    float tt, u, v;
                    int fr;
                    if (!tri_intersect(o, d, V3(tr->v0[0], tr->v0[1], tr->v0[2]),
                                       V3(tr->e1[0], tr->e1[1], tr->e1[2]),
                                       V3(tr->e2[0], tr->e2[1], tr->e2[2]), &tt, &u, &v, &fr))
                        continue;
                    if (!(tt > tmin)) continue;
                    uint32_t id = fbits(tr->v0[3]);
                    if (tt < best_t || (tt == best_t && best != NO_HIT && id < best)) {
                        best_t = tt; best = id; bu = u; bv = v; bfront = fr;
                    }
                }
            }
            /* internal children in the ray-octant order of the slots: slot
             * (k ^ oct) pushed for k = 7 .. 0, so slot oct is popped first */
            const float cull = best_t * BOX_PAD;
            for (int k = 7; k >= 0; --k) {
                int c = k ^ oct;
                if (!chit[c] || !(nd->meta[c] & 0x80)) continue;
                if (tnear[c] > cull) continue;
                st_node[sp] = nd->child_base + (nd->meta[c] & 0x7F);
                st_t[sp] = tnear[c];
                sp++;
            }
        }
        int hit = best != NO_HIT;
        if (hits) {
            hits[4 * r] = fbits(hit ? best_t : INFINITY);
            hits[4 * r + 1] = best;
            hits[4 * r + 2] = fbits(hit ? bu : 0.0f);
            hits[4 * r + 3] = fbits(hit ? bv : 0.0f);
        }
        if (flags) flags[r] = hit ? (1u | (bfront ? 2u : 0u)) : 0u;
        if (node_visits) node_visits[r] = nv;
        if (tri_tests) tri_tests[r] = nt;
    }
}

/* 30-bit Morton code of the triangle-AABB centroid over the centroid bounds. */
static inline uint32_t expand_bits10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
void or_morton_codes(const or_scene* s, uint32_t* codes) {
    float cmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, cmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float* c = (float*)malloc(sizeof(float) * 3 * (s->ntri + 1));
    for (uint32_t t = 0; t < s->ntri; ++t)
        for (int a = 0; a < 3; ++a) {
            float p0 = s->wv[9 * t + a], p1 = s->wv[9 * t + 3 + a], p2 = s->wv[9 * t + 6 + a];
            float lo = fminf(p0, fminf(p1, p2)), hi = fmaxf(p0, fmaxf(p1, p2));
            float cc = (lo + hi) * 0.5f;
            c[3 * t + a] = cc;
            if (cc < cmn[a]) cmn[a] = cc;
            if (cc > cmx[a]) cmx[a] = cc;
        }
    for (uint32_t t = 0; t < s->ntri; ++t) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            float ext = cmx[a] - cmn[a];
            float x = ext > 0.0f ? (c[3 * t + a] - cmn[a]) / ext : 0.0f;
            float y = x * 1024.0f;
            uint32_t qi = (uint32_t)y;
            q[a] = qi > 1023u ? 1023u : qi;
        }
        codes[t] = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
    }
    free(c);
}

/* ====================================================================== */
/* Shading helpers shared by photon and camera paths                        */
/* ====================================================================== */
/* shader.wgsl:530-545 + :573-579 world normal and front-face normal.       */
static v3 hit_world_normal(const or_scene* s, uint32_t tri, float u, float v) {
    uint32_t inst = s->tri_inst[tri], prim = s->tri_prim[tri];
    const or_instance* in = &s->inst[inst];
    const wsrt_mesh_info* mi = &s->meshes[in->mesh];
    uint32_t io = mi->index_offset + prim * 3u;
    const or_f32* n0 = s->verts[s->idx[io] + mi->vertex_offset].normal;
    const or_f32* n1 = s->verts[s->idx[io + 1] + mi->vertex_offset].normal;
    const or_f32* n2 = s->verts[s->idx[io + 2] + mi->vertex_offset].normal;
    float w = 1.0f - u - v;
    v3 ln = V3(n0[0] * w + n1[0] * u + n2[0] * v, n0[1] * w + n1[1] * u + n2[1] * v,
               n0[2] * w + n1[2] * u + n2[2] * v);
    ln = normalize3(ln);
    const float* W = in->w2o;
    v3 wn = V3(dot3(ln, V3(W[0], W[1], W[2])), dot3(ln, V3(W[3], W[4], W[5])),
               dot3(ln, V3(W[6], W[7], W[8])));
    return normalize3(wn);
}

static v3 random_unit_vector(uint32_t* rng) { /* shader.wgsl:162-169 */
    float u1 = or_rand(rng);
    float u2 = or_rand(rng);
    float z = 1.0f - 2.0f * u1;
    float r = sqrtf(maxf(0.0f, 1.0f - z * z));
    float phi = 2.0f * OR_PI * u2;
    return V3(r * cosf(phi), r * sinf(phi), z);
}

/* ====================================================================== */
/* Photon pass: photon_emit.wgsl:154-331                                     */
/* ====================================================================== */
/* Emission of photon path id (photon_emit.wgsl:159-208): the light, a point
 * on it, the cosine direction, the hero wavelength and the energy; returns
 * 0 when the scene has no light. */
static int photon_emit(const or_scene* s, uint32_t frame, uint32_t id, uint32_t* rng_io, v3* pos_out,
                       v3* dir_out, float* wl_out, float* energy_out) {
    uint32_t rng = or_pcg(id + frame * 927163u);
    uint32_t nl = s->nlights;
    if (nl == 0) return 0;
    uint32_t li = (uint32_t)(or_rand(&rng) * (float)nl);
    if (li > nl - 1) li = nl - 1;
    const wsrt_light_info* L = &s->lights[li];
    v3 emit_pos, emit_dir;
    float area_factor; /* light_area (4 for the point light's 4*pi) */
    if (L->params[0] == 1.0f) { /* point light (extension) */
        emit_pos = V3(L->v0[0], L->v0[1], L->v0[2]);
        emit_dir = random_unit_vector(&rng);
        area_factor = 4.0f;
    } else {
        v3 a = V3(L->v0[0], L->v0[1], L->v0[2]), b = V3(L->v1[0], L->v1[1], L->v1[2]),
           c = V3(L->v2[0], L->v2[1], L->v2[2]);
        v3 e1 = sub3(b, a), e2 = sub3(c, a);
        v3 nrm = normalize3(cross3(e1, e2));
        float area = length3(cross3(e1, e2)) * 0.5f;
        float r1 = or_rand(&rng), r2 = or_rand(&rng), sr = sqrtf(r1);
        float u = 1.0f - sr, v = r2 * sr, w = 1.0f - u - v;
        emit_pos = add3(add3(mul3s(a, u), mul3s(b, v)), mul3s(c, w));
        emit_dir = normalize3(add3(nrm, random_unit_vector(&rng)));
        area_factor = area;
    }
    float temp = L->params[1], intensity = L->params[2];
    float cmin, crange;
    or_cauchy_cdf(&cmin, &crange);
    float wl = or_cauchy_sample(or_rand(&rng), cmin, crange);
    float pdf = or_cauchy_pdf(wl, crange);
    float radiance = or_blackbody(wl, temp);
    float flux = (radiance / pdf) * area_factor * OR_PI;
    *energy_out = flux * intensity;
    *wl_out = wl;
    *pos_out = emit_pos;
    *dir_out = emit_dir;
    *rng_io = rng;
    return 1;
}

void or_photon_emit_f(const or_scene* s, uint32_t frame, uint32_t id, float* pos, float* dir, float* wl,
                      float* energy) {
    uint32_t rng;
    v3 p = V3(0, 0, 0), d = V3(0, 0, 0);
    *wl = 0.0f;
    *energy = 0.0f;
    photon_emit(s, frame, id, &rng, &p, &d, wl, energy);
    pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

static void trace_photon(const or_scene* s, const wsrt_config* cfg, uint32_t frame, uint32_t id,
                         wsrt_photon* out) {
    memset(out, 0, sizeof *out);
    uint32_t rng;
    v3 emit_pos, emit_dir;
    float wl, energy;
    if (!photon_emit(s, frame, id, &rng, &emit_pos, &emit_dir, &wl, &energy)) return;
    v3 pos = add3(emit_pos, mul3s(emit_dir, 0.001f));
    v3 dir = emit_dir;
    int caustic = 0;
    for (uint32_t depth = 0; depth < cfg->photon_max_depth; ++depth) {
        hit_t h = scene_closest(s, pos, dir, 0.0001f, 100.0f, 0);
        if (h.tri == NO_HIT) break;
        const wsrt_material* m = &s->mats[s->inst[s->tri_inst[h.tri]].mat];
        v3 wn = hit_world_normal(s, h.tri, h.u, h.v);
        v3 ff = h.front ? wn : neg3(wn);
        uint32_t type = (uint32_t)m->extra[0];
        v3 hp = add3(pos, mul3s(dir, h.t));
        if (type == 0u) {
            if (caustic) {
                out->position[0] = hp.x; out->position[1] = hp.y; out->position[2] = hp.z;
                out->direction[0] = -dir.x; out->direction[1] = -dir.y; out->direction[2] = -dir.z;
                out->wavelength = wl;
                out->energy = energy;
            }
            break;
        } else if (type == 1u) {
            dir = reflect3(normalize3(dir), ff);
            pos = add3(hp, mul3s(ff, 0.001f));
        } else if (type == 2u) {
            float A = m->extra[2], B = m->extra[3];
            float lum = wl / 1000.0f;
            float ir = A + (B / (lum * lum));
            float ratio = h.front ? 1.0f / ir : ir;
            v3 ud = normalize3(dir);
            float ct = minf(dot3(neg3(ud), ff), 1.0f);
            float st = sqrtf(1.0f - ct * ct);
            float r0 = (1.0f - ratio) / (1.0f + ratio);
            r0 = r0 * r0;
            float refl = r0 + (1.0f - r0) * powf(1.0f - ct, 5.0f);
            if (ratio * st > 1.0f || refl > or_rand(&rng)) {
                dir = reflect3(ud, ff);
            } else {
                caustic = 1;
                dir = refract3(ud, ff, ratio);
            }
            v3 off = dot3(dir, ff) > 0.0f ? ff : neg3(ff);
            pos = add3(hp, mul3s(off, 0.001f));
        } else if (type == 4u) {
            float n1 = 1.0f, n2 = m->extra[2], n3 = m->extra[3];
            float d = maxf(m->extra[1], 50.0f);
            v3 ud = normalize3(dir);
            float ct = minf(dot3(neg3(ud), ff), 1.0f);
            float ratio = h.front ? n3 / n1 : n1 / n3;
            float stsq = ratio * ratio * (1.0f - ct * ct);
            if (stsq > 1.0f) {
                dir = reflect3(ud, ff);
                caustic = 1;
            } else {
                float F = or_thin_film(ct, n1, n2, n3, d, wl);
                if (or_rand(&rng) < F) dir = reflect3(ud, ff);
                else dir = refract3(ud, ff, ratio);
                caustic = 1;
            }
            /* ray origin is not advanced in this branch (photon_emit.wgsl:291-326) */
        } else {
            break;
        }
    }
}

uint32_t or_photon_trace(const or_scene* s, const wsrt_config* cfg, uint32_t frame, uint32_t first,
                         uint32_t n, wsrt_photon* out) {
    uint32_t stored = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : stored)
    for (int64_t k = 0; k < (int64_t)n; ++k) {
        trace_photon(s, cfg, frame, first + (uint32_t)k, &out[k]);
        stored += out[k].energy != 0.0f;
    }
    return stored;
}

/* ====================================================================== */
/* Photon grid: build_grid.wgsl:25-73                                        */
/* ====================================================================== */
uint32_t or_grid_counting(const wsrt_photon* ph, uint32_t n, uint32_t frame, float cell,
                          uint32_t hs, uint32_t* cell_id, uint32_t* count, uint32_t* offset,
                          uint32_t* sorted_slot) {
    v3 j = or_grid_jitter(frame, cell);
    memset(count, 0, sizeof(uint32_t) * hs);
    uint32_t valid = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (ph[i].energy == 0.0f) { cell_id[i] = 0xFFFFFFFFu; continue; }
        uint32_t c = or_hash_cell(V3(ph[i].position[0], ph[i].position[1], ph[i].position[2]), j, cell, hs);
        cell_id[i] = c;
        count[c]++;
        valid++;
    }
    uint32_t run = 0;
    for (uint32_t c = 0; c < hs; ++c) { offset[c] = run; run += count[c]; }
    uint32_t* cursor = (uint32_t*)calloc(hs, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t c = cell_id[i];
        if (c == 0xFFFFFFFFu) continue;
        sorted_slot[offset[c] + cursor[c]++] = i;
    }
    free(cursor);
    return valid;
}

void or_grid_linked(const wsrt_photon* ph, uint32_t n, uint32_t frame, float cell, uint32_t hs,
                    uint32_t* head, uint32_t* next) {
    v3 j = or_grid_jitter(frame, cell);
    for (uint32_t c = 0; c < hs; ++c) head[c] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < n; ++i) {
        next[i] = 0xFFFFFFFFu;
        if (ph[i].energy == 0.0f) continue;
        uint32_t h = or_hash_cell(V3(ph[i].position[0], ph[i].position[1], ph[i].position[2]), j, cell, hs);
        next[i] = head[h];
        head[h] = i;
    }
}

/* ====================================================================== */
/* Camera paths: shader.wgsl:297-823                                         */
/* ====================================================================== */
struct or_frame {
    const or_scene* s;
    wsrt_config cfg;
    wsrt_camera_uniform cam;
    uint32_t W, H;
    wsrt_photon* photons;
    uint32_t* cell_id; uint32_t* count; uint32_t* offset; uint32_t* sorted;
    uint32_t stored;
    float M[16]; /* view_inv * proj_inv */
};

/* shader.wgsl:297-342, gathering the first gather_cap photons of the cell
 * in slot order (the counting-sort order). */
static v4 gather(const or_frame* f, v3 hp, v3 n, v4 wl) {
    v4 total = V4s(0.0f);
    const wsrt_config* c = &f->cfg;
    v3 j = or_grid_jitter(f->cam.frame_count, c->cell_size);
    uint32_t h = or_hash_cell(hp, j, c->cell_size, c->hash_size);
    float r = c->gather_radius * powf((float)(f->cam.frame_count + 1u), c->gather_radius_exponent);
    float r2 = r * r;
    uint32_t cnt = f->count[h];
    if (cnt > c->gather_cap) cnt = c->gather_cap;
    uint32_t off = f->offset[h];
    for (uint32_t k = 0; k < cnt; ++k) {
        const wsrt_photon* p = &f->photons[f->sorted[off + k]];
        v3 dd = sub3(hp, V3(p->position[0], p->position[1], p->position[2]));
        float dsq = dot3(dd, dd);
        if (dsq < r2) {
            if (dot3(n, V3(p->direction[0], p->direction[1], p->direction[2])) > 0.0f) {
                float dist = sqrtf(dsq);
                float sw = 1.0f - (dist / r);
                float es = p->energy * sw;
                float sr = c->spectral_radius;
                v4 spw = V4(maxf(1.0f - fabsf(wl.x - p->wavelength) / sr, 0.0f),
                            maxf(1.0f - fabsf(wl.y - p->wavelength) / sr, 0.0f),
                            maxf(1.0f - fabsf(wl.z - p->wavelength) / sr, 0.0f),
                            maxf(1.0f - fabsf(wl.w - p->wavelength) / sr, 0.0f));
                total = add4(total, mul4s(spw, es));
            }
        }
    }
    float spatial = 3.0f / (OR_PI * r2);
    float spectral = 1.0f / c->spectral_radius;
    float countn = 1.0f / (float)c->photon_norm;
    float density = spatial * spectral * countn;
    return mul4s(total, density);
}

typedef struct { v3 dir; float dist; float pdf; v4 radiance; int delta; } light_sample;

/* shader.wgsl:477-528 (+ point lights as an extension) */
static light_sample sample_light(const or_scene* s, v3 hp, v4 wl, uint32_t* rng) {
    light_sample ls;
    memset(&ls, 0, sizeof ls);
    uint32_t nl = s->nlights;
    if (nl == 0) return ls;
    uint32_t li = (uint32_t)(or_rand(rng) * (float)nl);
    if (li > nl - 1) li = nl - 1;
    const wsrt_light_info* L = &s->lights[li];
    if (L->params[0] == 1.0f) {
        (void)or_rand(rng);
        (void)or_rand(rng);
        v3 dl = sub3(V3(L->v0[0], L->v0[1], L->v0[2]), hp);
        float dsq = dot3(dl, dl);
        float dist = sqrtf(dsq);
        ls.dir = mul3s(dl, 1.0f / dist);
        ls.dist = dist;
        ls.pdf = dsq / (float)nl;
        ls.delta = 1;
        ls.radiance = mul4s(or_blackbody4(wl, L->params[1]), L->params[2]);
        return ls;
    }
    float r1 = or_rand(rng), r2 = or_rand(rng), sr = sqrtf(r1);
    float u = 1.0f - sr, v = r2 * sr, w = 1.0f - u - v;
    v3 a = V3(L->v0[0], L->v0[1], L->v0[2]), b = V3(L->v1[0], L->v1[1], L->v1[2]),
       c = V3(L->v2[0], L->v2[1], L->v2[2]);
    v3 p = add3(add3(mul3s(a, u), mul3s(b, v)), mul3s(c, w));
    v3 e1 = sub3(b, a), e2 = sub3(c, a);
    v3 nrm = normalize3(cross3(e1, e2));
    float area = length3(cross3(e1, e2)) * 0.5f;
    v3 dl = sub3(p, hp);
    float dsq = dot3(dl, dl);
    float dist = sqrtf(dsq);
    v3 l = V3(dl.x / dist, dl.y / dist, dl.z / dist);
    float cl = maxf(dot3(nrm, neg3(l)), 0.0f);
    if (cl <= 0.0001f) return ls;
    ls.dir = l;
    ls.dist = dist;
    ls.pdf = dsq / (area * (float)nl * cl);
    ls.radiance = mul4s(or_blackbody4(wl, L->params[1]), L->params[2]);
    return ls;
}

/* shader.wgsl:171-188, WGSL sign(0) = 0 */
static void onb(v3 n, v3* right, v3* up, int sign_fix) {
    float sg = n.z > 0.0f ? 1.0f : (n.z < 0.0f ? -1.0f : (sign_fix ? 1.0f : 0.0f));
    float a = -1.0f / (sg + n.z);
    float b = n.x * n.y * a;
    *right = V3(1.0f + sg * n.x * n.x * a, sg * b, -sg * n.x);
    *up = V3(b, sg + n.y * n.y * a, -n.y);
}

static float g1_schlick(float ndv, float r) { /* shader.wgsl:207-213 */
    float k = (r * r) / 2.0f;
    return ndv / (ndv * (1.0f - k) + k);
}

typedef struct { v3 dir; v4 att; int absorbed; } scatter_t;

/* shader.wgsl:347-471 */
static scatter_t sample_bsdf(v3 rdir, v3 ff, const wsrt_material* m, int front, v4 wl, v4 tp,
                             uint32_t* rng, uint32_t quirks) {
    scatter_t rec;
    rec.absorbed = 0;
    rec.dir = V3(0, 0, 0);
    rec.att = V4s(0.0f);
    uint32_t type = (uint32_t)m->extra[0];
    v3 v = normalize3(neg3(rdir));
    v4 spec = or_rgb_to_spectrum4(V3(m->color[0], m->color[1], m->color[2]), wl);
    if (type == 1u || type == 5u) {
        /* 5 (extension): GGX with a thin-film Fresnel factor (film color[0] nm,
         * index extra[2], substrate extra[3]) at the half-vector angle */
        float rough = maxf(m->extra[1], 0.001f);
        float a2 = rough * rough;
        float r1 = or_rand(rng), r2 = or_rand(rng);
        float phi = 2.0f * OR_PI * r1;
        float ct = sqrtf((1.0f - r2) / (1.0f + (a2 - 1.0f) * r2));
        float st = sqrtf(maxf(0.0f, 1.0f - ct * ct));
        v3 ht = V3(cosf(phi) * st, sinf(phi) * st, ct);
        v3 rt, up;
        onb(ff, &rt, &up, (quirks & WSRT_FIX_ONB_SIGN) != 0);
        v3 h = normalize3(add3(add3(mul3s(rt, ht.x), mul3s(up, ht.y)), mul3s(ff, ht.z)));
        v3 l = reflect3(neg3(v), h);
        float ndl = dot3(ff, l), ndv = dot3(ff, v);
        if (ndl > 0.0f && ndv > 0.0f) {
            float ndh = maxf(dot3(ff, h), 0.0f), vdh = maxf(dot3(v, h), 0.0f);
            v4 F;
            if (type == 5u) {
                float d = maxf(m->color[0], 50.0f);
                F = V4(or_thin_film(vdh, 1.0f, m->extra[2], m->extra[3], d, wl.x),
                       or_thin_film(vdh, 1.0f, m->extra[2], m->extra[3], d, wl.y),
                       or_thin_film(vdh, 1.0f, m->extra[2], m->extra[3], d, wl.z),
                       or_thin_film(vdh, 1.0f, m->extra[2], m->extra[3], d, wl.w));
            } else {
                float p5 = powf(clampf(1.0f - vdh, 0.0f, 1.0f), 5.0f);
                F = V4(spec.x + (1.0f - spec.x) * p5, spec.y + (1.0f - spec.y) * p5,
                       spec.z + (1.0f - spec.z) * p5, spec.w + (1.0f - spec.w) * p5);
            }
            float G = g1_schlick(ndl, rough) * g1_schlick(ndv, rough);
            float den = maxf(ndv * ndh, 0.0001f);
            rec.dir = l;
            rec.att = div4s(mul4s(mul4s(F, G), vdh), den);
        } else {
            rec.absorbed = 1;
        }
    } else if (type == 2u) {
        float A = m->extra[2], B = m->extra[3];
        float lum = wl.x / 1000.0f;
        float ir = A + (B / (lum * lum));
        float ratio = front ? 1.0f / ir : ir;
        v3 ud = normalize3(rdir);
        float ct = minf(dot3(neg3(ud), ff), 1.0f);
        float si2 = maxf(0.0f, 1.0f - ct * ct);
        float st2 = ratio * ratio * si2;
        if (st2 >= 1.0f) {
            rec.dir = reflect3(ud, ff);
            rec.att = V4s(1.0f);
        } else {
            float ctt = sqrtf(1.0f - st2);
            float cf = front ? ct : ctt;
            if (or_reflectance(cf, ratio) > or_rand(rng)) {
                rec.dir = reflect3(ud, ff);
                rec.att = V4s(1.0f);
            } else {
                rec.dir = refract3(ud, ff, ratio);
                float scale = ratio * ratio;
                v4 base = mul4s(spec, scale);
                if (B > 0.0f) {
                    int collapsed = tp.y < 0.0001f && tp.z < 0.0001f && tp.w < 0.0001f;
                    float mult = collapsed ? 1.0f : 4.0f;
                    rec.att = V4(base.x * mult, 0.0f, 0.0f, 0.0f);
                } else {
                    rec.att = base;
                }
            }
        }
    } else if (type == 4u) {
        float n2 = m->extra[2], n3 = m->extra[3], d = maxf(m->extra[1], 50.0f);
        v3 ud = normalize3(rdir);
        float c1 = minf(dot3(neg3(ud), ff), 1.0f);
        v4 F = V4(or_thin_film(c1, 1.0f, n2, n3, d, wl.x), or_thin_film(c1, 1.0f, n2, n3, d, wl.y),
                  or_thin_film(c1, 1.0f, n2, n3, d, wl.z), or_thin_film(c1, 1.0f, n2, n3, d, wl.w));
        float p = F.x;
        if (or_rand(rng) < p) {
            rec.dir = reflect3(ud, ff);
            rec.att = div4s(F, p);
        } else {
            rec.dir = ud;
            v4 T = V4(1.0f - F.x, 1.0f - F.y, 1.0f - F.z, 1.0f - F.w);
            rec.att = div4s(T, 1.0f - p);
        }
    } else {
        v3 sd = add3(ff, random_unit_vector(rng));
        if (length3(sd) < 0.001f) sd = ff;
        rec.dir = normalize3(sd);
        rec.att = spec;
    }
    return rec;
}

void or_scene_set_environment(or_scene* s, const float* rgb, uint32_t w, uint32_t h) {
    free(s->env);
    s->env = malloc(sizeof(float) * 3 * (size_t)w * h);
    memcpy(s->env, rgb, sizeof(float) * 3 * (size_t)w * h);
    s->env_w = w;
    s->env_h = h;
}

/* environment radiance of a miss: nearest texel of u = atan2(z, x)/2pi + 1/2,
 * v = acos(y)/pi; spectrum = max(rgb) * rgb_to_spectrum(rgb / max(rgb)) */
static v4 env_radiance(const or_scene* s, v3 d, v4 wl) {
    float u = atan2f(d.z, d.x) * (0.5f / OR_PI) + 0.5f;
    float v = acosf(clampf(d.y, -1.0f, 1.0f)) * (1.0f / OR_PI);
    float fx = maxf(u * (float)s->env_w, 0.0f), fy = maxf(v * (float)s->env_h, 0.0f);
    uint32_t x = (uint32_t)fx, y = (uint32_t)fy;
    if (x > s->env_w - 1u) x = s->env_w - 1u;
    if (y > s->env_h - 1u) y = s->env_h - 1u;
    const float* t = s->env + 3 * ((size_t)y * s->env_w + x);
    float mx = maxf(maxf(t[0], t[1]), t[2]);
    if (!(mx > 0.0f)) return V4s(0.0f);
    float im = 1.0f / mx;
    return mul4s(or_rgb_to_spectrum4(V3(t[0] * im, t[1] * im, t[2] * im), wl), mx);
}

/* Russian roulette (shader.wgsl:696-700): survival probability max(tp),
 * not clamped to 1 (so a throughput above 1 always survives and is divided
 * down); returns 0 when the path ends. */
static int roulette(v4* tp, uint32_t* rng, uint32_t quirks) {
    float ps = maxf(maxf(tp->x, tp->y), maxf(tp->z, tp->w));
    if (quirks & WSRT_FIX_RR_CLAMP) ps = minf(ps, 1.0f);
    if (or_rand(rng) > ps) return 0;
    *tp = div4s(*tp, ps);
    return 1;
}

int or_roulette_f(const float* tp4, uint32_t rng, uint32_t quirks, float* tp_out4, uint32_t* rng_out) {
    v4 tp = V4(tp4[0], tp4[1], tp4[2], tp4[3]);
    int alive = roulette(&tp, &rng, quirks);
    tp_out4[0] = tp.x; tp_out4[1] = tp.y; tp_out4[2] = tp.z; tp_out4[3] = tp.w;
    *rng_out = rng;
    return alive;
}

/* shader.wgsl:547-703 */
static v4 ray_color(const or_frame* f, v3 o, v3 d, v4 wl, uint32_t* rng) {
    const or_scene* s = f->s;
    const wsrt_config* c = &f->cfg;
    v4 acc = V4s(0.0f), tp = V4s(1.0f);
    int has_diffuse = 0, caustic = 0;
    float prev_pdf = 0.0f; /* cosine pdf of the last Lambert scattering; 0 after a specular one */
    for (uint32_t i = 0; i < c->max_depth; ++i) {
        hit_t h = scene_closest(s, o, d, 0.0001f, 100.0f, 0);
        if (h.tri == NO_HIT) {
            if (c->env_map && s->env) acc = add4(acc, mul4(tp, env_radiance(s, d, wl)));
            break;
        }
        const wsrt_material* m = &s->mats[s->inst[s->tri_inst[h.tri]].mat];
        v3 wn = hit_world_normal(s, h.tri, h.u, h.v);
        v3 ff = h.front ? wn : neg3(wn);
        uint32_t type = (uint32_t)m->extra[0];
        if (type == 3u) {
            if (caustic) break;
            v4 em = mul4s(or_blackbody4(wl, m->color[0]), m->color[1]);
            if (h.front || m->extra[1] > 0.0f) {
                if ((c->quirks & WSRT_FIX_EMISSION_MIS) && prev_pdf > 0.0f) {
                    /* shader.wgsl:601-626's balance weight, applied, with the
                     * world-space area of the hit triangle */
                    v3 e1 = V3(s->e1[3 * h.tri], s->e1[3 * h.tri + 1], s->e1[3 * h.tri + 2]);
                    v3 e2 = V3(s->e2[3 * h.tri], s->e2[3 * h.tri + 1], s->e2[3 * h.tri + 2]);
                    float area = length3(cross3(e1, e2)) * 0.5f;
                    float cl = maxf(dot3(ff, neg3(d)), 0.0001f);
                    float lp = (h.t * h.t) / (area * cl * (float)s->nlights);
                    acc = add4(acc, mul4s(mul4(em, tp), prev_pdf / (prev_pdf + lp)));
                } else {
                    acc = add4(acc, mul4(em, tp));
                }
            }
            break;
        }
        if (type == 0u) {
            has_diffuse = 1;
            caustic = 0;
        } else if (type == 2u) {
            if (has_diffuse) caustic = 1;
        }
        v3 hp = add3(o, mul3s(d, h.t));
        if (type == 0u) {
            v4 spec = or_rgb_to_spectrum4(V3(m->color[0], m->color[1], m->color[2]), wl);
            if (c->enable_photons) {
                v4 g = gather(f, hp, ff, wl);
                acc = add4(acc, mul4(mul4(tp, div4s(spec, OR_PI)), g));
            }
            light_sample ls = sample_light(s, hp, wl, rng);
            if (ls.pdf > 0.0f) {
                if (!scene_any(s, hp, ls.dir, 0.001f, ls.dist - 0.001f, 0)) {
                    float ct = maxf(dot3(ff, ls.dir), 0.0f);
                    v4 be = div4s(spec, OR_PI);
                    float bp = ct / OR_PI;
                    float w = ls.delta ? 1.0f / ls.pdf : (ls.pdf / (ls.pdf + bp)) / ls.pdf;
                    acc = add4(acc, mul4s(mul4s(mul4(mul4(tp, be), ls.radiance), ct), w));
                }
            }
        }
        scatter_t sc = sample_bsdf(d, ff, m, h.front, wl, tp, rng, c->quirks);
        if (sc.absorbed) break;
        prev_pdf = type == 0u ? maxf(dot3(ff, sc.dir), 0.0001f) / OR_PI : 0.0f;
        tp = mul4(tp, sc.att);
        o = hp;
        d = sc.dir;
        if (i > c->rr_after && !roulette(&tp, rng, c->quirks)) break;
    }
    return acc;
}

static void frame_matrix(const wsrt_camera_uniform* cam, float* M) {
    float vi[16], pi[16];
    for (int i = 0; i < 16; ++i) {
        vi[i] = cam->view_inverse[i];
        pi[i] = cam->proj_inverse[i];
    }
    or_mat4_mul(vi, pi, M);
}

/* shader.wgsl:709-799: one camera path; returns rgb, wavelengths, radiance. */
static v3 trace_pixel(const or_frame* f, uint32_t x, uint32_t y, v4* wl_out, v4* L_out) {
    uint32_t rng = or_pcg(x + y * f->W + f->cam.frame_count * 927163u);
    const float* M = f->M;
    v3 origin = V3(f->cam.view_inverse[12], f->cam.view_inverse[13], f->cam.view_inverse[14]);
    float jx = or_rand(&rng), jy = or_rand(&rng);
    float ux = ((float)x + jx) / (float)f->W, uy = ((float)y + jy) / (float)f->H;
    float nx = ux * 2.0f - 1.0f, ny = uy * 2.0f - 1.0f;
    float v[4] = {nx, ny, 1.0f, 1.0f}, tw[4];
    mat4_mul_vec4(M, v, tw);
    v3 dir = normalize3(sub3(V3(tw[0] / tw[3], tw[1] / tw[3], tw[2] / tw[3]), origin));
    float u = or_rand(&rng);
    float rot = or_rand(&rng);
    v4 wl = hero_wavelengths(u, rot);
    v4 L = ray_color(f, origin, dir, wl, &rng);
    v3 xyz = V3(0, 0, 0);
    v3 c;
    c = or_cie_xyz(wl.x); xyz = add3(xyz, mul3s(c, L.x));
    c = or_cie_xyz(wl.y); xyz = add3(xyz, mul3s(c, L.y));
    c = or_cie_xyz(wl.z); xyz = add3(xyz, mul3s(c, L.z));
    c = or_cie_xyz(wl.w); xyz = add3(xyz, mul3s(c, L.w));
    float cmin, crange;
    or_cauchy_cdf(&cmin, &crange);
    float sp = or_cauchy_pdf(wl.x, crange) + or_cauchy_pdf(wl.y, crange) +
               or_cauchy_pdf(wl.z, crange) + or_cauchy_pdf(wl.w, crange);
    xyz = mul3s(xyz, 1.0f / sp);
    *wl_out = wl;
    *L_out = L;
    return or_xyz_to_srgb(xyz);
}

or_frame* or_frame_create(const or_scene* s, const wsrt_config* cfg, const wsrt_camera_uniform* cam,
                          uint32_t W, uint32_t H) {
    or_frame* f = (or_frame*)calloc(1, sizeof(or_frame));
    f->s = s;
    f->cfg = *cfg;
    f->cam = *cam;
    f->W = W;
    f->H = H;
    frame_matrix(cam, f->M);
    uint32_t np = cfg->photons_per_pass;
    f->photons = (wsrt_photon*)calloc(np ? np : 1, sizeof(wsrt_photon));
    f->cell_id = (uint32_t*)malloc(sizeof(uint32_t) * (np ? np : 1));
    f->sorted = (uint32_t*)malloc(sizeof(uint32_t) * (np ? np : 1));
    f->count = (uint32_t*)calloc(cfg->hash_size, sizeof(uint32_t));
    f->offset = (uint32_t*)calloc(cfg->hash_size, sizeof(uint32_t));
    if (cfg->enable_photons && np) {
        or_photon_trace(s, cfg, cam->frame_count, 0, np, f->photons);
        f->stored = or_grid_counting(f->photons, np, cam->frame_count, cfg->cell_size, cfg->hash_size,
                                     f->cell_id, f->count, f->offset, f->sorted);
    }
    return f;
}
/* Same frame, with the photon buffer (slot order) given instead of traced,
 * so camera-path parity can be checked independently of the photon pass. */
or_frame* or_frame_create_with_photons(const or_scene* s, const wsrt_config* cfg,
                                       const wsrt_camera_uniform* cam, uint32_t W, uint32_t H,
                                       const wsrt_photon* photons, uint32_t n) {
    wsrt_config c = *cfg;
    c.enable_photons = 0;
    or_frame* f = or_frame_create(s, &c, cam, W, H);
    f->cfg.enable_photons = cfg->enable_photons;
    uint32_t np = cfg->photons_per_pass;
    if (n > np) n = np;
    if (cfg->enable_photons && np) {
        memcpy(f->photons, photons, sizeof(wsrt_photon) * n);
        f->stored = or_grid_counting(f->photons, np, cam->frame_count, cfg->cell_size, cfg->hash_size,
                                     f->cell_id, f->count, f->offset, f->sorted);
    }
    return f;
}

void or_frame_destroy(or_frame* f) {
    if (!f) return;
    free(f->photons); free(f->cell_id); free(f->sorted); free(f->count); free(f->offset);
    free(f);
}
uint32_t or_frame_photons_stored(const or_frame* f) { return f->stored; }

void or_frame_trace_pixels(const or_frame* f, const uint32_t* px, uint32_t n, float* wl,
                           float* rad, float* rgb) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        uint32_t p = px[i];
        v4 w, L;
        v3 c = trace_pixel(f, p % f->W, p / f->W, &w, &L);
        if (wl) { wl[4 * i] = w.x; wl[4 * i + 1] = w.y; wl[4 * i + 2] = w.z; wl[4 * i + 3] = w.w; }
        if (rad) { rad[4 * i] = L.x; rad[4 * i + 1] = L.y; rad[4 * i + 2] = L.z; rad[4 * i + 3] = L.w; }
        if (rgb) { rgb[3 * i] = c.x; rgb[3 * i + 1] = c.y; rgb[3 * i + 2] = c.z; }
    }
}

/* shader.wgsl:801-822 */
/* shader.wgsl:809-822: mean, clamp at 0, Reinhard on luma, RGBA8 without
 * gamma. */
static void tone_map(const float* na, uint8_t* out) {
    v3 fc = V3(maxf(na[0] / na[3] * 1.0f, 0.0f), maxf(na[1] / na[3] * 1.0f, 0.0f), maxf(na[2] / na[3] * 1.0f, 0.0f));
    float luma = dot3(fc, V3(0.2126f, 0.7152f, 0.0722f));
    float ml = luma / (luma + 1.0f);
    if (luma > 0.00001f) fc = mul3s(fc, ml / luma);
    float ch[4] = {fc.x, fc.y, fc.z, 1.0f};
    for (int k = 0; k < 4; ++k) out[k] = (uint8_t)lrintf(clampf(ch[k], 0.0f, 1.0f) * 255.0f);
}

void or_frame_render(const or_frame* f, float* accum, uint8_t* rgba8) {
    int64_t npx = (int64_t)f->W * f->H;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < npx; ++p) {
        v4 w, L;
        v3 c = trace_pixel(f, (uint32_t)(p % f->W), (uint32_t)(p / f->W), &w, &L);
        float* a = accum + 4 * p;
        v4 cur = f->cam.frame_count > 0 ? V4(a[0], a[1], a[2], a[3]) : V4s(0.0f);
        v4 na = add4(cur, V4(c.x, c.y, c.z, 1.0f));
        a[0] = na.x; a[1] = na.y; a[2] = na.z; a[3] = na.w;
        if (rgba8) tone_map(a, rgba8 + 4 * p);
    }
}

void or_tone_map_f(const float* accum, uint32_t n, uint8_t* rgba8) {
    for (uint32_t p = 0; p < n; ++p) tone_map(accum + 4 * p, rgba8 + 4 * p);
}

void or_frame_gather(const or_frame* f, const float* p, const float* n, const float* wl, float* out) {
    v4 g = gather(f, V3(p[0], p[1], p[2]), V3(n[0], n[1], n[2]), V4(wl[0], wl[1], wl[2], wl[3]));
    out[0] = g.x; out[1] = g.y; out[2] = g.z; out[3] = g.w;
}

/* cornell_box.rs:6-110 */
void or_scene_cornell_box(or_scene* s) {
    const float PI_F = 3.14159265358979323846f, HALF_PI = 1.57079632679489661923f;
    wsrt_material m_light = {{6500.0f, 1.0f, 0.0f, 1.0f}, {3, 0, 0, 0}};
    wsrt_material m_red = {{0.65f, 0.05f, 0.05f, 1}, {0, 0, 0, 0}};
    wsrt_material m_green = {{0.12f, 0.45f, 0.15f, 1}, {0, 0, 0, 0}};
    wsrt_material m_white = {{0.73f, 0.73f, 0.73f, 1}, {0, 0, 0, 0}};
    wsrt_material m_glass = {{1, 1, 1, 1}, {2, 0, 1.5f, 0.01f}};
    wsrt_material m_bubble = {{1, 1, 1, 1}, {4, 500.0f, 1.33f, 1.0f}};
    uint32_t light = or_scene_add_material(s, &m_light);
    uint32_t red = or_scene_add_material(s, &m_red);
    uint32_t green = or_scene_add_material(s, &m_green);
    uint32_t white = or_scene_add_material(s, &m_white);
    uint32_t glass = or_scene_add_material(s, &m_glass);
    uint32_t bubble = or_scene_add_material(s, &m_bubble);
    wsrt_vertex pv[4], cv[24];
    uint32_t pi[6], ci[36], nv, ni;
    or_geometry_plane(pv, pi, &nv, &ni);
    or_geometry_cube(cv, ci, &nv, &ni);
    or_geometry_sphere(4, NULL, NULL, &nv, &ni);
    wsrt_vertex* sv = (wsrt_vertex*)malloc(sizeof(wsrt_vertex) * nv);
    uint32_t* si = (uint32_t*)malloc(sizeof(uint32_t) * ni);
    or_geometry_sphere(4, sv, si, &nv, &ni);
    uint32_t plane = or_scene_add_mesh(s, pv, 4, pi, 6);
    uint32_t cube = or_scene_add_mesh(s, cv, 24, ci, 36);
    uint32_t sphere = or_scene_add_mesh(s, sv, nv, si, ni);
    free(sv);
    free(si);
    float T[16], R[16], S[16], TR[16], M[16];
#define TRS(tx, ty, tz, rot_fn, ang, sx, sy, sz)                                 \
    do {                                                                         \
        or_mat4_translation(tx, ty, tz, T);                                      \
        rot_fn(ang, R);                                                          \
        or_mat4_scale(sx, sy, sz, S);                                            \
        or_mat4_mul(T, R, TR);                                                   \
        or_mat4_mul(TR, S, M);                                                   \
    } while (0)
    /* floor: translation * scale */
    or_mat4_translation(0, -1, 0, T);
    or_mat4_scale(2, 2, 2, S);
    or_mat4_mul(T, S, M);
    or_scene_add_instance(s, plane, white, M);
    TRS(0, 1, 0, or_mat4_rotation_x, PI_F, 2, 2, 2);
    or_scene_add_instance(s, plane, white, M);
    TRS(0, 0, -1, or_mat4_rotation_x, HALF_PI, 2, 2, 2);
    or_scene_add_instance(s, plane, white, M);
    TRS(-1, 0, 0, or_mat4_rotation_z, -HALF_PI, 2, 2, 2);
    or_scene_add_instance(s, plane, red, M);
    TRS(1, 0, 0, or_mat4_rotation_z, HALF_PI, 2, 2, 2);
    or_scene_add_instance(s, plane, green, M);
    TRS(0, 0.99f, 0, or_mat4_rotation_x, PI_F, 0.5f, 0.5f, 0.5f);
    const float lp[4] = {0.0f, 6500.0f, 1.0f, 0.0f};
    or_scene_add_light_instance(s, plane, light, M, lp);
    TRS(-0.35f, -0.4f + 0.002f, -0.3f, or_mat4_rotation_y, 0.4f, 0.6f, 1.2f, 0.6f);
    or_scene_add_instance(s, cube, glass, M);
    or_mat4_translation(0.4f, -0.64f, 0.3f, T);
    or_mat4_scale(0.75f, 0.75f, 0.75f, S);
    or_mat4_mul(T, S, M);
    or_scene_add_instance(s, sphere, bubble, M);
#undef TRS
}

/* ====================================================================== */
/* Taps of single steps, for the pins of the oracle's transport code        */
/* ====================================================================== */
void or_unit_vectors_f(uint32_t seed, uint32_t n, float* out) {
    uint32_t rng = or_pcg(seed);
    for (uint32_t i = 0; i < n; ++i) {
        v3 u = random_unit_vector(&rng);
        out[3 * i] = u.x; out[3 * i + 1] = u.y; out[3 * i + 2] = u.z;
    }
}

int or_sample_bsdf_f(const float* rdir, const float* ffn, const wsrt_material* m, int front, const float* wl4,
                     const float* tp4, uint32_t rng, uint32_t quirks, float* dir_out, float* att_out,
                     uint32_t* rng_out) {
    scatter_t sc = sample_bsdf(V3(rdir[0], rdir[1], rdir[2]), V3(ffn[0], ffn[1], ffn[2]), m, front,
                               V4(wl4[0], wl4[1], wl4[2], wl4[3]), V4(tp4[0], tp4[1], tp4[2], tp4[3]), &rng,
                               quirks);
    dir_out[0] = sc.dir.x; dir_out[1] = sc.dir.y; dir_out[2] = sc.dir.z;
    att_out[0] = sc.att.x; att_out[1] = sc.att.y; att_out[2] = sc.att.z; att_out[3] = sc.att.w;
    *rng_out = rng;
    return sc.absorbed;
}

int or_sample_light_f(const or_scene* s, const float* hp, const float* wl4, uint32_t rng, float* dir,
                      float* dist, float* pdf, float* rad4, uint32_t* rng_out) {
    light_sample ls = sample_light(s, V3(hp[0], hp[1], hp[2]), V4(wl4[0], wl4[1], wl4[2], wl4[3]), &rng);
    dir[0] = ls.dir.x; dir[1] = ls.dir.y; dir[2] = ls.dir.z;
    *dist = ls.dist;
    *pdf = ls.pdf;
    rad4[0] = ls.radiance.x; rad4[1] = ls.radiance.y; rad4[2] = ls.radiance.z; rad4[3] = ls.radiance.w;
    *rng_out = rng;
    return ls.delta;
}

void or_ray_color_f(const or_frame* f, const float* o, const float* d, const float* wl4, uint32_t rng,
                    float* out4) {
    v4 L = ray_color(f, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), V4(wl4[0], wl4[1], wl4[2], wl4[3]), &rng);
    out4[0] = L.x; out4[1] = L.y; out4[2] = L.z; out4[3] = L.w;
}
