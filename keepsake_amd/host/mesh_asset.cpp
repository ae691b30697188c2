// Mesh assets: OBJ reader, binary mesh asset reader/writer, [mesh_asset.*] parser and the
// convert_mesh_asset task (mesh_asset.cpp:70-310).
#include "mesh_asset.h"

#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <tuple>

namespace ks {

void MeshData::transform(const std::array<float, 16> &m) {
    // inverse transpose of the upper 3x3 (for normals), in double
    double a[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = m[4 * r + c];
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (det == 0.0) throw ConfigError("mesh to_world transform is singular");
    // (A^-1)^T = cofactor matrix / det
    double it[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            it[r][c] = (a[r1][c1] * a[r2][c2] - a[r1][c2] * a[r2][c1]) / det;
        }
    for (size_t i = 0; i < vertex_count(); ++i) {
        float *v = &vertices[3 * i];
        float x = v[0], y = v[1], z = v[2];
        for (int r = 0; r < 3; ++r) v[r] = m[4 * r] * x + m[4 * r + 1] * y + m[4 * r + 2] * z + m[4 * r + 3];
        if (!vertex_normals.empty()) {
            float *n = &vertex_normals[3 * i];
            double o[3];
            for (int r = 0; r < 3; ++r) o[r] = it[r][0] * n[0] + it[r][1] * n[1] + it[r][2] * n[2];
            double l = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
            for (int r = 0; r < 3; ++r) n[r] = (float)(l > 0.0 ? o[r] / l : o[r]);
        }
    }
}

namespace {

struct ObjIndex {
    int v, vt, vn;
    bool operator<(const ObjIndex &o) const { return std::tie(v, vt, vn) < std::tie(o.v, o.vt, o.vn); }
};

int resolve_index(int i, size_t count, int line, const fs::path &path) {
    // OBJ indices are 1-based; negative ones count back from the end of the list read so far
    int r = i > 0 ? i - 1 : (int)count + i;
    if (i == 0 || r < 0 || (size_t)r >= count)
        throw ConfigError(path.string() + ":" + std::to_string(line) + ": index " + std::to_string(i) + " out of range");
    return r;
}


// Corner triples of one polygon face, as the OBJ reader's triangulation makes them: triangles as
// written; quads split along the shorter diagonal ([0,1,2],[0,2,3] when |v2-v0| < |v3-v1|, else
// [0,1,3],[1,2,3]; tinyobjloader 2.x ObjReader with triangulate = true); larger polygons by ear
// clipping in the plane that drops the largest component of the Newell normal (convex corner with
// no other corner inside), falling back to a fan over what remains when no ear is found.
std::vector<int> triangulate_polygon(const std::vector<float> &pos, const std::vector<int> &v) {
    const size_t n = v.size();
    std::vector<int> out;
    auto P = [&](int c, int k) { return pos[3 * (size_t)v[(size_t)c] + k]; };
    if (n == 3) return {0, 1, 2};
    if (n == 4) {
        float d02 = 0, d13 = 0;
        for (int k = 0; k < 3; ++k) {
            float a = P(2, k) - P(0, k), b = P(3, k) - P(1, k);
            d02 += a * a;
            d13 += b * b;
        }
        if (d02 < d13) return {0, 1, 2, 0, 2, 3};
        return {0, 1, 3, 1, 2, 3};
    }
    double nrm[3] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        size_t j = (i + 1) % n;
        for (int k = 0; k < 3; ++k) {
            int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
            nrm[k] += ((double)P((int)i, k1) - P((int)j, k1)) * ((double)P((int)i, k2) + P((int)j, k2));
        }
    }
    int drop = 0;
    for (int k = 1; k < 3; ++k)
        if (std::fabs(nrm[k]) > std::fabs(nrm[drop])) drop = k;
    int ax = (drop + 1) % 3, ay = (drop + 2) % 3;
    double sign = nrm[drop] >= 0 ? 1.0 : -1.0;
    std::vector<int> rem(n);
    for (size_t i = 0; i < n; ++i) rem[i] = (int)i;
    auto cross2 = [&](int a, int b, int c) {
        double x0 = P(b, ax) - P(a, ax), y0 = P(b, ay) - P(a, ay);
        double x1 = P(c, ax) - P(a, ax), y1 = P(c, ay) - P(a, ay);
        return sign * (x0 * y1 - y0 * x1);
    };
    while (rem.size() > 3) {
        size_t m = rem.size();
        bool clipped = false;
        for (size_t i = 0; i < m && !clipped; ++i) {
            int a = rem[(i + m - 1) % m], b = rem[i], c = rem[(i + 1) % m];
            if (cross2(a, b, c) <= 0.0) continue; // reflex or degenerate corner
            bool inside = false;
            for (size_t j = 0; j < m && !inside; ++j) {
                int q = rem[j];
                if (q == a || q == b || q == c) continue;
                inside = cross2(a, b, q) >= 0.0 && cross2(b, c, q) >= 0.0 && cross2(c, a, q) >= 0.0;
            }
            if (inside) continue;
            out.insert(out.end(), {a, b, c});
            rem.erase(rem.begin() + (long)i);
            clipped = true;
        }
        if (!clipped) break;
    }
    for (size_t k = 1; k + 1 < rem.size(); ++k) out.insert(out.end(), {rem[0], rem[k], rem[k + 1]});
    return out;
}
} // namespace

MeshAsset load_obj(const fs::path &path, bool twosided, bool use_smooth_normal) {
    std::ifstream f(path);
    if (!f) throw ConfigError("cannot open mesh asset " + path.string());
    std::vector<float> pos, tex, nrm;
    MeshAsset asset;
    MeshData cur;
    std::map<ObjIndex, uint32_t> remap;
    std::string name;
    auto finish = [&]() {
        if (!cur.indices.empty()) {
            cur.name = name;
            cur.twosided = twosided;
            cur.use_smooth_normal = use_smooth_normal;
            asset.meshes.push_back(std::move(cur));
        }
        cur = MeshData();
        remap.clear();
    };
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "v") { // missing coordinates read as 0; trailing vertex colours are ignored
            float x = 0, y = 0, z = 0;
            if (!(ls >> x)) throw ConfigError(path.string() + ":" + std::to_string(lineno) + ": bad vertex");
            ls >> y >> z;
            pos.insert(pos.end(), {x, y, z});
        } else if (tag == "vt") {
            float u = 0, v = 0;
            ls >> u >> v;
            tex.insert(tex.end(), {u, v});
        } else if (tag == "vn") {
            float x = 0, y = 0, z = 0;
            if (!(ls >> x)) throw ConfigError(path.string() + ":" + std::to_string(lineno) + ": bad normal");
            ls >> y >> z;
            nrm.insert(nrm.end(), {x, y, z});
        } else if (tag == "o" || tag == "g") {
            finish();
            std::string rest;
            std::getline(ls, rest);
            size_t b = rest.find_first_not_of(" \t");
            name = b == std::string::npos ? std::string() : rest.substr(b);
        } else if (tag == "f") {
            std::vector<ObjIndex> corners;
            std::string tok;
            while (ls >> tok) {
                ObjIndex idx{0, -1, -1};
                int parts[3] = {0, 0, 0};
                bool has[3] = {false, false, false};
                size_t s0 = 0;
                for (int k = 0; k < 3 && s0 <= tok.size(); ++k) {
                    size_t e = tok.find('/', s0);
                    std::string field = tok.substr(s0, e == std::string::npos ? std::string::npos : e - s0);
                    if (!field.empty()) {
                        parts[k] = std::stoi(field);
                        has[k] = true;
                    }
                    if (e == std::string::npos) break;
                    s0 = e + 1;
                }
                if (!has[0]) throw ConfigError(path.string() + ":" + std::to_string(lineno) + ": face without vertex");
                idx.v = resolve_index(parts[0], pos.size() / 3, lineno, path);
                if (has[1]) idx.vt = resolve_index(parts[1], tex.size() / 2, lineno, path);
                if (has[2]) idx.vn = resolve_index(parts[2], nrm.size() / 3, lineno, path);
                corners.push_back(idx);
            }
            if (corners.size() < 3)
                throw ConfigError(path.string() + ":" + std::to_string(lineno) + ": face with < 3 vertices");
            std::vector<int> v_of(corners.size());
            for (size_t k = 0; k < corners.size(); ++k) v_of[k] = corners[k].v;
            // triangulate on the OBJ corners, then remap corners in triangle order (first use numbers a vertex)
            for (int c : triangulate_polygon(pos, v_of)) {
                const ObjIndex &idx = corners[(size_t)c];
                auto ins = remap.insert({idx, (uint32_t)cur.vertex_count()});
                if (ins.second) {
                    cur.vertices.insert(cur.vertices.end(), pos.begin() + 3 * idx.v, pos.begin() + 3 * idx.v + 3);
                    // "Either all or none vertices have uv / vertex normal" (mesh_asset.cpp:128-146)
                    if (idx.vt >= 0) {
                        if (cur.vertex_count() > 1 && cur.texcoords.empty())
                            throw ConfigError(path.string() + ": either all or no vertices have uv");
                        cur.texcoords.insert(cur.texcoords.end(), tex.begin() + 2 * idx.vt, tex.begin() + 2 * idx.vt + 2);
                    } else if (!cur.texcoords.empty()) {
                        throw ConfigError(path.string() + ": either all or no vertices have uv");
                    }
                    if (idx.vn >= 0) {
                        if (cur.vertex_count() > 1 && cur.vertex_normals.empty())
                            throw ConfigError(path.string() + ": either all or no vertices have a vertex normal");
                        cur.vertex_normals.insert(cur.vertex_normals.end(), nrm.begin() + 3 * idx.vn,
                                                  nrm.begin() + 3 * idx.vn + 3);
                    } else if (!cur.vertex_normals.empty()) {
                        throw ConfigError(path.string() + ": either all or no vertices have a vertex normal");
                    }
                }
                cur.indices.push_back(ins.first->second);
            }
        }
        // mtllib, usemtl, s, l, p: not used by this renderer
    }
    finish();
    return asset;
}

namespace {
constexpr const char *kBinaryMagic = "i_am_a_binary_mesh_asset";

struct Reader {
    std::ifstream s;
    fs::path path;
    template <typename T> T get() {
        T v;
        s.read(reinterpret_cast<char *>(&v), sizeof(T));
        if (!s) throw ConfigError("truncated binary mesh asset " + path.string());
        return v;
    }
    template <typename T> void get_array(std::vector<T> &v, size_t n) {
        v.resize(n);
        s.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(sizeof(T) * n));
        if (!s) throw ConfigError("truncated binary mesh asset " + path.string());
    }
};
} // namespace

MeshAsset load_mesh_binary(const fs::path &path) {
    Reader r{std::ifstream(path, std::ios::binary), path};
    if (!r.s) throw ConfigError("cannot open mesh asset " + path.string());
    char magic[32] = {0};
    r.s.read(magic, (std::streamsize)std::strlen(kBinaryMagic));
    if (!r.s || std::strcmp(magic, kBinaryMagic) != 0) throw ConfigError("invalid binary mesh asset " + path.string());
    int n_meshes = r.get<int>();
    if (n_meshes < 0) throw ConfigError("invalid binary mesh asset " + path.string());
    MeshAsset a;
    a.meshes.resize((size_t)n_meshes);
    for (MeshData &m : a.meshes) {
        m.twosided = r.get<bool>();
        m.use_smooth_normal = r.get<bool>();
        int nv = r.get<int>(), nt = r.get<int>();
        bool has_tc = r.get<bool>(), has_vn = r.get<bool>();
        if (nv < 0 || nt < 0) throw ConfigError("invalid binary mesh asset " + path.string());
        r.get_array(m.vertices, 3 * (size_t)nv + 1); // + 1 Embree padding float on disk
        m.vertices.resize(3 * (size_t)nv);
        r.get_array(m.indices, 3 * (size_t)nt);
        for (uint32_t i : m.indices)
            if (i >= (uint32_t)nv) throw ConfigError("binary mesh asset index out of range in " + path.string());
        if (has_tc) {
            r.get_array(m.texcoords, 2 * (size_t)nv + 2);
            m.texcoords.resize(2 * (size_t)nv);
        }
        if (has_vn) {
            r.get_array(m.vertex_normals, 3 * (size_t)nv + 1);
            m.vertex_normals.resize(3 * (size_t)nv);
        }
    }
    return a;
}

void write_mesh_binary(const MeshAsset &asset, const fs::path &path) {
    std::ofstream s(path, std::ios::binary);
    if (!s) throw ConfigError("cannot write " + path.string());
    auto put = [&](const void *p, size_t n) { s.write(reinterpret_cast<const char *>(p), (std::streamsize)n); };
    put(kBinaryMagic, std::strlen(kBinaryMagic));
    int n = (int)asset.meshes.size();
    put(&n, sizeof(int));
    const float pad = std::nanf("");
    for (const MeshData &m : asset.meshes) {
        bool ts = m.twosided, sn = m.use_smooth_normal, tc = !m.texcoords.empty(), vn = !m.vertex_normals.empty();
        int nv = (int)m.vertex_count(), nt = (int)m.tri_count();
        put(&ts, 1);
        put(&sn, 1);
        put(&nv, sizeof(int));
        put(&nt, sizeof(int));
        put(&tc, 1);
        put(&vn, 1);
        put(m.vertices.data(), sizeof(float) * m.vertices.size());
        put(&pad, sizeof(float));
        put(m.indices.data(), sizeof(uint32_t) * m.indices.size());
        if (tc) {
            put(m.texcoords.data(), sizeof(float) * m.texcoords.size());
            put(&pad, sizeof(float));
            put(&pad, sizeof(float));
        }
        if (vn) {
            put(m.vertex_normals.data(), sizeof(float) * m.vertex_normals.size());
            put(&pad, sizeof(float));
        }
    }
    if (!s) throw ConfigError("failed writing " + path.string());
}

MeshAsset create_mesh_asset(const ConfigArgs &args) {
    fs::path path = args.load_path("path");
    std::string fmt = args.load_string("format", "obj");
    MeshAsset a;
    if (fmt == "obj") {
        if (args.load_bool("load_materials"))
            throw ConfigError("load_materials = true (MTL materials) is not supported; assign materials with the "
                              "task's material list");
        a = load_obj(path, args.load_bool("twosided", false), args.load_bool("use_smooth_normal", true));
    } else if (fmt == "bin") {
        a = load_mesh_binary(path);
    } else {
        throw ConfigError("Unsupported mesh asset format [" + fmt + "]");
    }
    if (args.contains("to_world")) {
        std::array<float, 16> m = args.load_transform("to_world");
        bool identity = true;
        for (int i = 0; i < 16; ++i) identity &= m[i] == ((i % 5 == 0) ? 1.0f : 0.0f);
        if (!identity)
            for (MeshData &md : a.meshes) md.transform(m);
    }
    return a;
}

void convert_mesh_asset(const ConfigArgs &args, const fs::path &task_dir, int task_id) {
    (void)task_id;
    ConfigArgs list = args["assets"];
    for (size_t i = 0; i < list.array_size(); ++i) {
        std::string asset_path = list[(int)i].node->str;
        MeshAsset a = create_mesh_asset(args.asset(asset_path));
        std::string name = asset_path.substr(asset_path.rfind('.') + 1);
        write_mesh_binary(a, task_dir / (name + ".bin"));
        std::printf("wrote %s (%zu meshes)\n", (task_dir / (name + ".bin")).string().c_str(), a.meshes.size());
    }
}

} // namespace ks
