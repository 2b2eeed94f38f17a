// ref_glm.cpp -- test infrastructure: InitTransformMatrix's composition
// (src/Transform.h:36-60) evaluated by the reference tree's own glm
// (external/glm), compiled by oracle/Makefile into oracle/_ref/ when the
// reference tree is present.  Checks or_xform.c and the library's parser.
#include <cstring>

#include <glm/glm.hpp>
#include <glm/gtc/matrix_transform.hpp>

extern "C" void ref_glm_transform(const float *t, const float *s, const float *r, const float *parent, float *out,
                                  float *inv)
{
    glm::vec3 tv(t[0], t[1], t[2]), sv(s[0], s[1], s[2]), rv(r[0], r[1], r[2]);
    glm::mat4 m = glm::translate(glm::mat4(1.0f), tv) *
                  glm::rotate(glm::mat4(1.0f), glm::radians(rv.z), glm::vec3(0.0f, 0.0f, 1.0f)) *
                  glm::rotate(glm::mat4(1.0f), glm::radians(rv.y), glm::vec3(0.0f, 1.0f, 0.0f)) *
                  glm::rotate(glm::mat4(1.0f), glm::radians(rv.x), glm::vec3(1.0f, 0.0f, 0.0f)) *
                  glm::scale(glm::mat4(1.0f), sv);
    if (parent) {
        glm::mat4 p;
        std::memcpy(&p[0][0], parent, 64);
        m = p * m;
    }
    glm::mat4 i = glm::inverse(m);
    std::memcpy(out, &m[0][0], 64);
    std::memcpy(inv, &i[0][0], 64);
}
