// obj_ingest.cpp -- OBJ ingest straight into scene-description arrays (rt_obj_load).
//
// Restates OBJModelImporter::Import (src/importer.cpp:29-146) without its
// per-triangle heap objects: the file is read once into memory and parsed into
// flat arrays in World::triangles order, ready for rt_scene_desc.
//
// Semantics kept from the reference:
//   - identifiers are whitespace-separated tokens (`fs >> identifier`); values
//     of o / v / vn / vt / s are read as further tokens of the stream, and an
//     unknown identifier skips the rest of its line;
//   - vertex, normal and uv lists start with a zero entry, so indices are
//     1-based; an index that is absent, unparsable or out of range gives zero;
//   - an `f` line is trimmed, runs of spaces collapse to one, and each
//     space-separated vertex splits on '/' into vertex / uv / normal indices
//     (std::stoi: leading integer, trailing characters ignored);
//   - polygons fan into triangles (v0, v_i, v_i+1) (Polygon::ConvertToTriangle,
//     src/object.cpp:333-342);
//   - `s` is read as a bool (num_get: an integer that must be 0 or 1), and a
//     numeric read that fails -- `s off`, `v` with a non-number -- ends the
//     import as the reference's stream loop does, keeping what was read;
//   - a mesh is listed when the next `o` starts or the file ends.
// Differences: a missing file is an error here (the reference silently
// imports nothing), and faces or `s` before any `o` go to an unnamed mesh
// (the reference dereferences a null mesh there).
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/rendertoy_amd.h"

namespace rtamd {
void set_error(const std::string &msg);
}

namespace {

struct Obj {
    rt_obj pub{};
    std::vector<float> vertices, normals;
    std::vector<uint32_t> tri_mesh;
    std::vector<uint8_t> smooth;
    std::vector<std::string> names;
    std::vector<const char *> name_ptrs;
};

struct Reader {
    const char *p, *end;
    bool failed = false;

    static bool ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
    void skip_ws() {
        while (p < end && ws(*p)) ++p;
    }
    bool token(std::string &out) {
        skip_ws();
        if (p >= end) return false;
        const char *b = p;
        while (p < end && !ws(*p)) ++p;
        out.assign(b, p);
        return true;
    }
    void skip_line() {
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
    std::string rest_of_line() {   // std::getline: up to, not including, '\n'
        const char *b = p;
        while (p < end && *p != '\n') ++p;
        std::string s(b, p);
        if (p < end) ++p;
        return s;
    }
    // istream >> float: skip whitespace, parse the longest number prefix.
    float number() {
        skip_ws();
        if (p >= end || !(std::isdigit((unsigned char)*p) || *p == '-' || *p == '+' || *p == '.')) {
            failed = true;
            return 0.0f;
        }
        std::string tok;
        const char *b = p;
        while (p < end && (std::isdigit((unsigned char)*p) || *p == '-' || *p == '+' || *p == '.' || *p == 'e' ||
                           *p == 'E'))
            ++p;
        tok.assign(b, p);
        char *e = nullptr;
        errno = 0;
        const float v = std::strtof(tok.c_str(), &e);
        if (e == tok.c_str()) {
            failed = true;
            return 0.0f;
        }
        p = b + (e - tok.c_str());   // characters strtof did not take stay in the stream
        return v;
    }
    // istream >> bool (noboolalpha): an integer; 0 or 1, else failure.
    bool boolean() {
        skip_ws();
        const char *b = p;
        if (p < end && (*p == '-' || *p == '+')) ++p;
        const char *digits = p;
        while (p < end && std::isdigit((unsigned char)*p)) ++p;
        if (p == digits) {
            failed = true;
            return false;
        }
        const long v = std::strtol(std::string(b, p).c_str(), nullptr, 10);
        if (v != 0 && v != 1) {
            failed = true;
            return true;
        }
        return v == 1;
    }
};

// std::stoi on a vertex-index field: leading whitespace, optional sign,
// digits; anything else is no value.
bool parse_index(const std::string &s, int &out) {
    size_t i = 0;
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    size_t j = i;
    if (j < s.size() && (s[j] == '-' || s[j] == '+')) ++j;
    size_t k = j;
    while (k < s.size() && std::isdigit((unsigned char)s[k])) ++k;
    if (k == j) return false;
    errno = 0;
    const long v = std::strtol(s.substr(i, k - i).c_str(), nullptr, 10);
    if (errno == ERANGE || v > 2147483647L || v < -2147483647L - 1) return false;
    out = (int)v;
    return true;
}

// Split (importer.cpp:15-25): std::getline pieces, no trailing empty piece.
std::vector<std::string> split(const std::string &s, char d) {
    std::vector<std::string> out;
    size_t b = 0;
    while (b < s.size()) {
        size_t e = s.find(d, b);
        if (e == std::string::npos) e = s.size();
        out.push_back(s.substr(b, e - b));
        b = e + 1;
    }
    return out;
}

// regex_replace("^ +| +$|( ) +", "$1"): trim spaces, collapse runs of spaces.
std::string tidy(const std::string &s) {
    std::string out;
    size_t b = 0, e = s.size();
    while (b < e && s[b] == ' ') ++b;
    while (e > b && s[e - 1] == ' ') --e;
    for (size_t i = b; i < e; ++i) {
        if (s[i] == ' ' && !out.empty() && out.back() == ' ') continue;
        out += s[i];
    }
    return out;
}

struct V3 {
    float x, y, z;
};

}  // namespace

extern "C" {

int rt_obj_load(const char *path, rt_obj **out) {
    if (!path || !out) {
        rtamd::set_error("path and out are required");
        return RT_ERR_INVALID;
    }
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        rtamd::set_error(std::string("cannot open ") + path);
        return RT_ERR_INVALID;
    }
    std::string data;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n > 0) {
        data.resize((size_t)n);
        if (std::fread(&data[0], 1, data.size(), f) != data.size()) data.clear();
    }
    std::fclose(f);
    Obj *o = new (std::nothrow) Obj();
    if (!o) {
        rtamd::set_error("host out of memory");
        return RT_ERR_OOM;
    }
    try {
        std::vector<V3> vert{{0, 0, 0}}, norm{{0, 0, 0}};
        int cur = -1;   // index of the mesh being filled (listed when finished)
        std::vector<std::pair<std::string, uint8_t>> meshes;
        auto start_mesh = [&](const std::string &name) {
            meshes.push_back({name, 0});
            cur = (int)meshes.size() - 1;
        };
        Reader r{data.data(), data.data() + data.size()};
        std::string id;
        while (!r.failed && r.token(id)) {
            if (id == "o") {
                std::string name;
                r.token(name);
                start_mesh(name);
            } else if (id == "v" || id == "vn") {
                V3 v;
                v.x = r.number();
                v.y = r.failed ? 0.0f : r.number();
                v.z = r.failed ? 0.0f : r.number();
                if (r.failed) break;
                (id == "v" ? vert : norm).push_back(v);
            } else if (id == "vt") {
                r.number();   // uvs are read past; the renderers do not use them
                if (!r.failed) r.number();
                if (r.failed) break;
            } else if (id == "s") {
                if (cur < 0) start_mesh("");
                meshes[cur].second = r.boolean() ? 1 : 0;
            } else if (id == "f") {
                if (cur < 0) start_mesh("");
                const std::vector<std::string> defs = split(tidy(r.rest_of_line()), ' ');
                std::vector<V3> pv, pn;
                for (const std::string &d : defs) {
                    const std::vector<std::string> parts = split(d, '/');
                    int idx[3] = {0, 0, 0};
                    bool has[3] = {false, false, false};
                    for (size_t k = 0; k < parts.size() && k < 3; ++k) has[k] = parse_index(parts[k], idx[k]);
                    pv.push_back(has[0] && idx[0] >= 0 && (size_t)idx[0] < vert.size() ? vert[idx[0]] : V3{0, 0, 0});
                    pn.push_back(has[2] && idx[2] >= 0 && (size_t)idx[2] < norm.size() ? norm[idx[2]] : V3{0, 0, 0});
                }
                for (size_t b = 1, c = 2; c < pv.size(); ++b, ++c) {
                    const V3 tv[3] = {pv[0], pv[b], pv[c]}, tn[3] = {pn[0], pn[b], pn[c]};
                    for (int k = 0; k < 3; ++k) {
                        o->vertices.insert(o->vertices.end(), {tv[k].x, tv[k].y, tv[k].z});
                        o->normals.insert(o->normals.end(), {tn[k].x, tn[k].y, tn[k].z});
                    }
                    o->tri_mesh.push_back((uint32_t)cur);
                }
            } else {
                r.skip_line();
            }
        }
        for (auto &m : meshes) {
            o->names.push_back(m.first);
            o->smooth.push_back(m.second);
        }
        for (auto &s : o->names) o->name_ptrs.push_back(s.c_str());
        rt_obj &p = o->pub;
        p.internal = o;
        p.num_triangles = (uint32_t)o->tri_mesh.size();
        p.vertices = o->vertices.data();
        p.normals = o->normals.data();
        p.tri_mesh = o->tri_mesh.data();
        p.num_meshes = (uint32_t)meshes.size();
        p.mesh_smooth = o->smooth.data();
        p.mesh_names = o->name_ptrs.data();
    } catch (const std::bad_alloc &) {
        delete o;
        rtamd::set_error("host out of memory");
        return RT_ERR_OOM;
    }
    *out = &o->pub;
    return RT_OK;
}

int rt_obj_destroy(rt_obj *obj) {
    if (obj) delete static_cast<Obj *>(obj->internal);
    return RT_OK;
}

}  // extern "C"
