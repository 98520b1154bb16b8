// Host-side checks of the RenderToy C++ API (include/RenderToy/): reads commands
// on stdin and prints results as hex float bits, for tests/test_api_cpu.py.
//   M tx ty tz ex ey ez            -> 16 words: Translation(t) * RotationEulerXYZ(e)
//   S gate_x gate_y focal W H      -> 2 words: PrepareScreenSpace top, right
//   O path                         -> triangles, meshes, emitters, then per triangle
//                                     13 words (vert_w[0], edges, normal, area) and its mesh
//   G size sigma orient            -> UnitizedGaussianKernel<float, All/X/Y> weights, row-major
#include <RenderToy/rendertoy.h>

#include <cstdint>
#include <cstring>
#include <iostream>
#include <map>
#include <string>

using namespace RenderToy;

static void hex(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    std::cout << std::hex << u << std::dec << " ";
}

struct Probe : Renderer {
    Probe(RenderContext *rc) : Renderer(rc) {}
    void Render() override {}
    void Screen(Camera *c, float &t, float &r) { PrepareScreenSpace(c, t, r); }
};

int main() {
    std::string cmd;
    while (std::cin >> cmd) {
        if (cmd == "M") {
            float t[3], e[3];
            std::cin >> t[0] >> t[1] >> t[2] >> e[0] >> e[1] >> e[2];
            Matrix4x4f m = AffineTransformation::Translation(Vector3f(t[0], t[1], t[2])) *
                           AffineTransformation::RotationEulerXYZ(Vector3f(e[0], e[1], e[2]));
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) hex(m[i][j]);
            std::cout << "\n";
        } else if (cmd == "S") {
            float gx, gy, f;
            size_t W, H;
            std::cin >> gx >> gy >> f >> W >> H;
            Camera cam(Matrix4x4f::I, f, Vector2f(gx, gy), 0.1f, 1000.0f);
            RenderContext rc(nullptr, FormatSettings(SizeN(W, H), Vector2f(1.0f, 1.0f)));
            Probe p(&rc);
            float top, right;
            p.Screen(&cam, top, right);
            hex(top);
            hex(right);
            std::cout << "\n";
        } else if (cmd == "G") {
            size_t size;
            float sigma;
            int orient;
            std::cin >> size >> sigma >> orient;
            IConvolutionKernel<float> k;
            if (orient == 0) k = UnitizedGaussianKernel<float, Orientation::All>(size, sigma);
            else if (orient == 1) k = UnitizedGaussianKernel<float, Orientation::X>(size, sigma);
            else k = UnitizedGaussianKernel<float, Orientation::Y>(size, sigma);
            for (size_t i = 0; i < k.kernel_mat.row; ++i)
                for (size_t j = 0; j < k.kernel_mat.column; ++j) hex(k.kernel_mat[i][j]);
            std::cout << "\n";
        } else if (cmd == "O") {
            std::string path;
            std::cin >> path;
            World w;
            OBJModelImporter::Import(w, path);
            PrincipledBSDF light(Vector3f::White, Vector3f(1.0f, 1.0f, 1.0f)), plain(Vector3f::White);
            std::map<const Mesh *, int> id;
            for (size_t i = 0; i < w.meshes.size(); ++i) {
                w.meshes[i]->tex = (i % 2) ? &light : &plain;
                id[w.meshes[i]] = (int)i;
            }
            w.PrepareDirectLightSampling();
            std::map<const Triangle *, int> tid;
            for (size_t i = 0; i < w.triangles.size(); ++i) tid[w.triangles[i]] = (int)i;
            std::cout << w.triangles.size() << " " << w.meshes.size() << " " << w.CountEmitters() << "\n";
            for (const Triangle *t : w.EmissiveTriangles()) std::cout << tid[t] << " ";
            std::cout << "\n";
            for (const Triangle *t : w.triangles) {
                const auto &v = t->WorldVertices();
                Vector3f e1 = v[1] - v[0], e2 = v[2] - v[0];
                for (int k = 0; k < 3; ++k) hex(v[0][k]);
                for (int k = 0; k < 3; ++k) hex(e1[k]);
                for (int k = 0; k < 3; ++k) hex(e2[k]);
                Vector3f n = t->GeometricalNormalC();
                for (int k = 0; k < 3; ++k) hex(n[k]);
                hex(t->AreaC());
                std::cout << id[t->parent] << "\n";
            }
        }
    }
    return 0;
}
