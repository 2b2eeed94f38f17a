/* or_xform.c -- test infrastructure (the CPU oracle), not product code.
 *
 * The transform of a .ymtmp Transform block: InitTransformMatrix
 * (src/Transform.h:36-60) composes parent * T * Rz * Ry * Rx * S with glm's
 * translate / rotate / scale (external/glm/glm/gtc/matrix_transform.inl),
 * its mat4 product (detail/type_mat4x4.inl:595-613) and its 4x4 inverse
 * (detail/func_matrix.inl:297-347).  Restated here in glm's operation order
 * (column-major m[col][row]); checked bit for bit against glm itself built
 * from the reference tree (oracle/_ref) where that tree is present. */
#include <math.h>
#include <string.h>

typedef struct { float m[4][4]; } xm4;

static xm4 xm_id(void)
{
    xm4 r;
    memset(&r, 0, sizeof(r));
    for (int i = 0; i < 4; ++i) r.m[i][i] = 1.0f;
    return r;
}

static xm4 xm_mul(const xm4 *a, const xm4 *b)
{
    xm4 r;
    for (int col = 0; col < 4; ++col) {
        for (int row = 0; row < 4; ++row) {
            float s = a->m[0][row] * b->m[col][0];
            s = s + a->m[1][row] * b->m[col][1];
            s = s + a->m[2][row] * b->m[col][2];
            s = s + a->m[3][row] * b->m[col][3];
            r.m[col][row] = s;
        }
    }
    return r;
}

static xm4 xm_translate(const float *v)
{
    xm4 i = xm_id(), r = i;
    for (int row = 0; row < 4; ++row) {
        float s = i.m[0][row] * v[0];
        s = s + i.m[1][row] * v[1];
        s = s + i.m[2][row] * v[2];
        r.m[3][row] = s + i.m[3][row];
    }
    return r;
}

static xm4 xm_scale(const float *v)
{
    xm4 i = xm_id(), r;
    for (int row = 0; row < 4; ++row) {
        r.m[0][row] = i.m[0][row] * v[0];
        r.m[1][row] = i.m[1][row] * v[1];
        r.m[2][row] = i.m[2][row] * v[2];
        r.m[3][row] = i.m[3][row];
    }
    return r;
}

static xm4 xm_rotate(float angle, float x, float y, float z)
{
    float c = cosf(angle), s = sinf(angle);
    float len2 = x * x + y * y + z * z;
    float il = 1.0f / sqrtf(len2);
    float a[3] = {x * il, y * il, z * il};
    float omc = 1.0f - c;
    float t[3] = {omc * a[0], omc * a[1], omc * a[2]};
    float R[3][3] = {
        {c + t[0] * a[0], t[0] * a[1] + s * a[2], t[0] * a[2] - s * a[1]},
        {t[1] * a[0] - s * a[2], c + t[1] * a[1], t[1] * a[2] + s * a[0]},
        {t[2] * a[0] + s * a[1], t[2] * a[1] - s * a[0], c + t[2] * a[2]},
    };
    xm4 i = xm_id(), r;
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 4; ++row) {
            float v = i.m[0][row] * R[col][0];
            v = v + i.m[1][row] * R[col][1];
            v = v + i.m[2][row] * R[col][2];
            r.m[col][row] = v;
        }
    for (int row = 0; row < 4; ++row) r.m[3][row] = i.m[3][row];
    return r;
}

static xm4 xm_inverse(const xm4 *M)
{
    const float (*m)[4] = M->m;
    float k00 = m[2][2] * m[3][3] - m[3][2] * m[2][3], k02 = m[1][2] * m[3][3] - m[3][2] * m[1][3];
    float k03 = m[1][2] * m[2][3] - m[2][2] * m[1][3], k04 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
    float k06 = m[1][1] * m[3][3] - m[3][1] * m[1][3], k07 = m[1][1] * m[2][3] - m[2][1] * m[1][3];
    float k08 = m[2][1] * m[3][2] - m[3][1] * m[2][2], k10 = m[1][1] * m[3][2] - m[3][1] * m[1][2];
    float k11 = m[1][1] * m[2][2] - m[2][1] * m[1][2], k12 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
    float k14 = m[1][0] * m[3][3] - m[3][0] * m[1][3], k15 = m[1][0] * m[2][3] - m[2][0] * m[1][3];
    float k16 = m[2][0] * m[3][2] - m[3][0] * m[2][2], k18 = m[1][0] * m[3][2] - m[3][0] * m[1][2];
    float k19 = m[1][0] * m[2][2] - m[2][0] * m[1][2], k20 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
    float k22 = m[1][0] * m[3][1] - m[3][0] * m[1][1], k23 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
    float F[6][4] = {{k00, k00, k02, k03}, {k04, k04, k06, k07}, {k08, k08, k10, k11},
                     {k12, k12, k14, k15}, {k16, k16, k18, k19}, {k20, k20, k22, k23}};
    float V[4][4] = {{m[1][0], m[0][0], m[0][0], m[0][0]}, {m[1][1], m[0][1], m[0][1], m[0][1]},
                     {m[1][2], m[0][2], m[0][2], m[0][2]}, {m[1][3], m[0][3], m[0][3], m[0][3]}};
    xm4 r;
    for (int k = 0; k < 4; ++k) {
        float sa = (k & 1) ? -1.0f : 1.0f, sb = -sa;
        r.m[0][k] = (V[1][k] * F[0][k] - V[2][k] * F[1][k] + V[3][k] * F[2][k]) * sa;
        r.m[1][k] = (V[0][k] * F[0][k] - V[2][k] * F[3][k] + V[3][k] * F[4][k]) * sb;
        r.m[2][k] = (V[0][k] * F[1][k] - V[1][k] * F[3][k] + V[3][k] * F[5][k]) * sa;
        r.m[3][k] = (V[0][k] * F[2][k] - V[1][k] * F[4][k] + V[2][k] * F[5][k]) * sb;
    }
    float d0 = m[0][0] * r.m[0][0], d1 = m[0][1] * r.m[1][0], d2 = m[0][2] * r.m[2][0], d3 = m[0][3] * r.m[3][0];
    float det = (d0 + d1) + (d2 + d3);
    float inv = 1.0f / det;
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) r.m[c][k] = r.m[c][k] * inv;
    return r;
}

/* out / inv: column-major 16 floats; parent NULL for a root transform;
 * rotate in degrees (glm::radians: deg * 0.0174532925...f) */
void or_glm_transform(const float *translate, const float *scale, const float *rotate_deg, const float *parent,
                      float *out, float *inv)
{
    const float k = (float)0.01745329251994329576923690768489;
    xm4 T = xm_translate(translate);
    xm4 Rz = xm_rotate(rotate_deg[2] * k, 0.0f, 0.0f, 1.0f);
    xm4 Ry = xm_rotate(rotate_deg[1] * k, 0.0f, 1.0f, 0.0f);
    xm4 Rx = xm_rotate(rotate_deg[0] * k, 1.0f, 0.0f, 0.0f);
    xm4 S = xm_scale(scale);
    xm4 a = xm_mul(&T, &Rz);
    xm4 b = xm_mul(&a, &Ry);
    xm4 c = xm_mul(&b, &Rx);
    xm4 M = xm_mul(&c, &S);
    if (parent) {
        xm4 P;
        memcpy(P.m, parent, 64);
        M = xm_mul(&P, &M);
    }
    xm4 I = xm_inverse(&M);
    memcpy(out, M.m, 64);
    memcpy(inv, I.m, 64);
}
