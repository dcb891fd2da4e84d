// glm_probe.cpp — test driver for the reference's vendored GLM 1.0.0
// (src/glm, header-only, compiled from where it lies; never copied).  It
// evaluates the GLM calls the reference makes for the render path so the
// oracle's and libvrt's restatements can be checked bit for bit:
//   normalize / length / dot          (src/fshader.glsl:136,147,155,168,173,177 use
//                                      these GLSL built-ins; GLM's order is the
//                                      oracle's documented choice)
//   mat4 * vec4                       (src/fshader.glsl:137, src/render.cpp:401)
//   rotate, mat4 * mat4               (src/controls.cpp:136-142, src/render.cpp:400)
//   radians                           (src/render.cpp:400)
// Input: one operation per line, "<op> <hex word> ...", floats as IEEE bits.
// Output: one line per operation with the result words in hex.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "glm/glm.hpp"
#include "glm/gtc/matrix_transform.hpp"

static float f_of(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
static uint32_t u_of(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

static glm::mat4 mat_of(const std::vector<float>& a, size_t o) {
    glm::mat4 m;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) m[c][r] = a[o + 4 * c + r];
    return m;
}

static void put(const std::vector<float>& out) {
    for (size_t i = 0; i < out.size(); i++) std::printf(i ? " %08x" : "%08x", u_of(out[i]));
    std::printf("\n");
}

static std::vector<float> mat_vec(const glm::mat4& m) {
    std::vector<float> v;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) v.push_back(m[c][r]);
    return v;
}

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream ss(line);
        std::string op;
        ss >> op;
        std::vector<float> a;
        std::string w;
        while (ss >> w) a.push_back(f_of((uint32_t)std::stoul(w, nullptr, 16)));
        if (op == "normalize") {
            glm::vec3 r = glm::normalize(glm::vec3(a[0], a[1], a[2]));
            put({r.x, r.y, r.z});
        } else if (op == "length") {
            put({glm::length(glm::vec3(a[0], a[1], a[2]))});
        } else if (op == "dot") {
            put({glm::dot(glm::vec3(a[0], a[1], a[2]), glm::vec3(a[3], a[4], a[5]))});
        } else if (op == "matvec") {
            glm::vec4 r = mat_of(a, 0) * glm::vec4(a[16], a[17], a[18], a[19]);
            put({r.x, r.y, r.z, r.w});
        } else if (op == "rotate") {
            glm::mat4 r = glm::rotate(mat_of(a, 0), a[16], glm::vec3(a[17], a[18], a[19]));
            put(mat_vec(r));
        } else if (op == "matmul") {
            put(mat_vec(mat_of(a, 0) * mat_of(a, 16)));
        } else if (op == "camera") {  // src/controls.cpp:136-142
            glm::mat4 rotX = glm::mat4(1.0f), rotY = glm::mat4(1.0f);
            rotX = glm::rotate(rotX, a[0], glm::vec3(1, 0, 0));
            rotY = glm::rotate(rotY, a[1], glm::vec3(0, 1, 0));
            put(mat_vec(rotY * rotX));
        } else if (op == "orbit") {  // src/render.cpp:400-401
            glm::mat4 rot = glm::mat4(1.0f);
            rot = glm::rotate(rot, glm::radians(a[0]), glm::vec3(0, 0, 1));
            glm::vec3 p = rot * glm::vec4(glm::vec3(a[1], a[2], a[3]), 1);
            put({p.x, p.y, p.z});
        } else if (op == "primary") {  // src/fshader.glsl:136-137 in GLM's order
            glm::vec3 d = glm::normalize(glm::vec3(a[2] * a[0], a[3], a[1]));
            glm::vec3 r = glm::vec3(mat_of(a, 4) * glm::vec4(d, 0));
            put({r.x, r.y, r.z});
        } else {
            std::fprintf(stderr, "unknown op %s\n", op.c_str());
            return 2;
        }
    }
    return 0;
}
