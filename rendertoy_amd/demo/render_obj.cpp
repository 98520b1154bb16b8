// render_obj -- renders an OBJ scene through the RenderToy C++ API.
//
//   render_obj <scene.obj> <scene.cfg> <kind> <out.raw>
//
// scene.cfg (text): "W H spp max_depth seed", "focal gate_x gate_y near far",
// 16 floats of the camera's object-to-world matrix, then one line of 24
// material floats (PrincipledBSDF field order) per mesh of the OBJ.
// kind: pt | normal | depth | albedo | mask.  out.raw: W*H*3 float32, buffer order.
#include <RenderToy/rendertoy.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

using namespace RenderToy;

int main(int argc, char **argv) {
    if (argc != 5) {
        std::cerr << "usage: render_obj scene.obj scene.cfg kind out.raw\n";
        return 2;
    }
    World world;
    OBJModelImporter::Import(world, argv[1]);
    std::ifstream cfg(argv[2]);
    size_t W, H;
    int spp;
    unsigned max_depth;
    unsigned long long seed;
    float focal, gx, gy, nearc, farc;
    cfg >> W >> H >> spp >> max_depth >> seed >> focal >> gx >> gy >> nearc >> farc;
    Matrix4x4f M;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) cfg >> M[i][j];
    std::vector<std::unique_ptr<PrincipledBSDF>> mats;
    for (Mesh *m : world.meshes) {
        float f[24];
        for (float &x : f) cfg >> x;
        mats.emplace_back(new PrincipledBSDF(Vector3f(f[0], f[1], f[2]), Vector3f(f[3], f[4], f[5]), f[6], f[7], f[8],
                                             f[9], f[10], f[11], f[12], f[13], f[14], f[15], f[16], f[17],
                                             Vector3f(f[18], f[19], f[20])));
        m->tex = mats.back().get();
    }
    if (!cfg) {
        std::cerr << "bad config\n";
        return 2;
    }
    world.PrepareDirectLightSampling();
    world.cameras.push_back(Camera(Matrix4x4f::I, focal, Vector2f(gx, gy), nearc, farc));
    world.cameras[0].SetO2W(M);
    RenderContext rc(&world, FormatSettings(SizeN(W, H), Vector2f(1.0f, 1.0f)));
    std::string kind = argv[3];
    std::unique_ptr<Renderer> r;
    if (kind == "pt") r.reset(new PathTracingRenderer(&rc, spp, max_depth, seed));
    else if (kind == "normal") r.reset(new NormalRenderer(&rc));
    else if (kind == "depth") r.reset(new DepthBufferRenderer(&rc));
    else if (kind == "albedo") r.reset(new AlbedoRenderer(&rc));
    else if (kind == "mask") r.reset(new IntersectTestRenderer(&rc));
    else {
        std::cerr << "unknown kind\n";
        return 2;
    }
    r->Render();
    std::ofstream out(argv[4], std::ios::binary);
    out.write(reinterpret_cast<const char *>(rc.buffer), (std::streamsize)(W * H * sizeof(Vector3f)));
    std::cout << "rendered " << world.triangles.size() << " triangles, " << world.meshes.size() << " meshes, "
              << world.CountEmitters() << " emitters\n";
    return 0;
}
