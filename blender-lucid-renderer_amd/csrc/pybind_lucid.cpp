// pybind_lucid.cpp -- the `lucidrenderer` Python module (host side, C++).
//
// Mirrors the reference's render-session API, PYBIND11_MODULE(lucidrenderer)
// in LucidRenderer/cpp_renderer/src/pybind_module.cpp:38-543 and class
// PyRenderer in include/pybind_renderer.hpp:62-1746: same class and method
// names, arguments, defaults, return conventions (SUM output, bottom row
// first, alpha 1, zeros without scene or camera, bool from load_scene_json)
// and GIL release during rendering.  All rendering goes through the C ABI of
// liblucid_hip.so (lucid_hip.h); there is no CPU fallback: without a usable
// HIP device a render raises RuntimeError.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "diag_host.h"
#include "json_lite.h"
#include "lucid_hip.h"

namespace py = pybind11;
using lucid::json::Value;

namespace {

// Host copy of one mesh: the arrays a lucid_mesh_desc points into.
struct HostMesh {
    std::vector<float> vertices;
    std::vector<int32_t> indices;
    std::vector<float> tri_normals;
    std::vector<float> grid;
    std::vector<float> tri_uvs;
    // node tree: names live in `strings` (stable addresses), sockets per node
    std::deque<std::string> strings;
    std::vector<std::vector<lucid_node_socket>> node_inputs, node_outputs;
    std::vector<lucid_node> nodes;
    lucid_mesh_desc desc{};
};

struct HostScene {
    std::vector<std::unique_ptr<HostMesh>> meshes;
    std::vector<lucid_light_desc> lights;
    float env_color[3] = {0.05f, 0.05f, 0.05f};  // Environment defaults (core/scene.hpp:175-178)
    float env_strength = 1.0f;
};

float f(const Value& v) { return (float)lucid::json::num(v); }
float f_at(const Value& v, size_t i) { return (float)lucid::json::at(v, i); }

// Material Output "Surface" socket linked? (pybind_renderer.hpp:1386-1398
// with parseNodeTree, node_evaluator.cpp:82-149: a socket is linked when a
// "links" entry targets (node name, socket name); a tree without nodes is
// invalid and leaves the surface linked).  Node trees are otherwise not
// evaluated.
bool surface_linked(const Value& mat) {
    const Value* nt = mat.get("node_tree");
    if (!nt || !nt->is_object()) return true;
    const Value* nodes = nt->get("nodes");
    if (!nodes || nodes->kind != Value::Arr || nodes->arr.empty()) return true;
    const Value* links = nt->get("links");
    for (const Value& n : nodes->arr) {
        const Value* ty = n.get("type");
        if (!ty || ty->kind != Value::Str || ty->s != "ShaderNodeOutputMaterial") continue;
        std::string name;
        if (const Value* nm = n.get("name"); nm && nm->kind == Value::Str) name = nm->s;
        const Value* ins = n.get("inputs");
        if (ins && ins->kind == Value::Arr) {
            for (const Value& sock : ins->arr) {
                const Value* sn = sock.get("name");
                if (!(sn && sn->kind == Value::Str && sn->s == "Surface")) continue;
                if (!links || links->kind != Value::Arr) return false;
                for (const Value& l : links->arr) {
                    const Value* tn = l.get("to_node");
                    const Value* ts = l.get("to_socket");
                    if (tn && ts && tn->kind == Value::Str && ts->kind == Value::Str && tn->s == name &&
                        ts->s == "Surface")
                        return true;
                }
                return false;
            }
        }
        return false;
    }
    return true;
}

const char* keep(HostMesh& hm, const std::string& text) {
    hm.strings.push_back(text);
    return hm.strings.back().c_str();
}

std::string str_or_empty(const Value* v) { return (v && v->kind == Value::Str) ? v->s : std::string(); }

// parseSocketValue (node_evaluator.cpp:47-69): a number is FLOAT, a numeric
// 3-array VEC3, a 4-array VEC4; anything else carries no value the
// evaluator reads.
void socket_value(const Value* v, lucid_node_socket& s) {
    s.kind = LUCID_SOCK_NONE;
    if (!v) return;
    if (v->kind == Value::Num) {
        s.kind = LUCID_SOCK_FLOAT;
        s.value[0] = (float)v->n;
    } else if (v->kind == Value::NumArr && v->shape.size() == 1 && (v->flat.size() == 3 || v->flat.size() == 4)) {
        s.kind = v->flat.size() == 3 ? LUCID_SOCK_VEC3 : LUCID_SOCK_VEC4;
        for (size_t k = 0; k < v->flat.size(); ++k) s.value[k] = (float)v->flat[k];
    }
}

// parseNodeTree (node_evaluator.cpp:82-149): nodes in JSON order, each input
// socket linked from the (node, socket) of the last "links" entry that
// targets it; useNodes = "use_nodes" && a tree with nodes
// (pybind_renderer.hpp:1381-1384).
void parse_node_tree(const Value& mat, HostMesh& hm, lucid_mesh_desc& d) {
    const Value* nt = mat.get("node_tree");
    if (!nt || !nt->is_object()) return;
    const Value* nodes = nt->get("nodes");
    if (!nodes || nodes->kind != Value::Arr || nodes->arr.empty()) return;
    std::map<std::pair<std::string, std::string>, std::pair<std::string, std::string>> conn;
    if (const Value* links = nt->get("links"); links && links->kind == Value::Arr) {
        for (const Value& l : links->arr) {
            auto field = [&](const char* key) {
                const Value* v = l.get(key);
                if (!v || v->kind != Value::Str) throw std::runtime_error(std::string("node_tree link without ") + key);
                return v->s;
            };
            conn[{field("to_node"), field("to_socket")}] = {field("from_node"), field("from_socket")};
        }
    }
    for (const Value& nj : nodes->arr) {
        const std::string name = str_or_empty(nj.get("name"));
        lucid_node n{};
        n.name = keep(hm, name);
        n.type = keep(hm, str_or_empty(nj.get("type")));
        std::vector<lucid_node_socket> ins, outs;
        if (const Value* a = nj.get("inputs"); a && a->kind == Value::Arr) {
            for (const Value& sj : a->arr) {
                lucid_node_socket s{};
                const std::string sn = str_or_empty(sj.get("name"));
                s.name = keep(hm, sn);
                socket_value(sj.get("default_value"), s);
                auto it = conn.find({name, sn});
                if (it != conn.end()) {
                    s.linked_node = keep(hm, it->second.first);
                    s.linked_socket = keep(hm, it->second.second);
                }
                ins.push_back(s);
            }
        }
        if (const Value* a = nj.get("outputs"); a && a->kind == Value::Arr) {
            for (const Value& sj : a->arr) {
                lucid_node_socket s{};
                s.name = keep(hm, str_or_empty(sj.get("name")));
                socket_value(sj.get("default_value"), s);
                outs.push_back(s);
            }
        }
        hm.node_inputs.push_back(std::move(ins));
        hm.node_outputs.push_back(std::move(outs));
        hm.nodes.push_back(n);
    }
    for (size_t i = 0; i < hm.nodes.size(); ++i) {
        hm.nodes[i].inputs = hm.node_inputs[i].data();
        hm.nodes[i].num_inputs = (int32_t)hm.node_inputs[i].size();
        hm.nodes[i].outputs = hm.node_outputs[i].data();
        hm.nodes[i].num_outputs = (int32_t)hm.node_outputs[i].size();
    }
    d.nodes = hm.nodes.data();
    d.num_nodes = (int32_t)hm.nodes.size();
    const Value* un = mat.get("use_nodes");
    d.use_nodes = (un && lucid::json::boolean(*un)) ? 1 : 0;
}

// loadSceneFromJsonString (pybind_renderer.hpp:1258-1542) and
// parseNativeLights field reading (:1544-1746); unit conversions happen in
// the library (lucid_scene_create).
void parse_scene(const std::string& text, HostScene& sc) {
    Value j = lucid::json::parse(text);
    if (!j.is_object()) throw std::runtime_error("scene JSON must be an object");
    const Value* meshes = j.get("meshes");
    if (!meshes) return;  // the reference returns the empty scene here
    if (!meshes->is_array()) throw std::runtime_error("\"meshes\" must be an array");
    if (meshes->kind == Value::Arr) {
        for (const Value& mj : meshes->arr) {
            auto hm = std::make_unique<HostMesh>();
            lucid_mesh_desc& d = hm->desc;
            std::memset(&d, 0, sizeof(d));
            if (const Value* vs = mj.get("vertices")) {
                int64_t n = vs->size();
                hm->vertices.resize(3 * n);
                if (vs->kind == Value::NumArr) {
                    if (n > 0 && (vs->shape.size() < 2 || vs->shape[1] < 3 || !vs->regular))
                        throw std::runtime_error("vertices must be [x, y, z] triples");
                    int64_t stride = n > 0 ? (int64_t)vs->flat.size() / n : 3;
                    for (int64_t i = 0; i < n; ++i)
                        for (int k = 0; k < 3; ++k) hm->vertices[3 * i + k] = (float)vs->flat[i * stride + k];
                } else {
                    for (int64_t i = 0; i < n; ++i)
                        for (int k = 0; k < 3; ++k) hm->vertices[3 * i + k] = f_at(vs->arr[i], k);
                }
            }
            int64_t ntn = 0;
            if (const Value* tn = mj.get("triangle_normals")) {
                ntn = tn->size();
                hm->tri_normals.resize(9 * ntn);
                if (tn->kind == Value::NumArr) {
                    if (ntn > 0 && (tn->flat.size() != (size_t)(9 * ntn) || !tn->regular))
                        throw std::runtime_error("triangle_normals must be 3 x [x, y, z] per triangle");
                    for (int64_t i = 0; i < 9 * ntn; ++i) hm->tri_normals[i] = (float)tn->flat[i];
                } else {
                    for (int64_t i = 0; i < ntn; ++i) {
                        const Value& t3 = tn->arr[i];
                        if (t3.kind != Value::NumArr || t3.flat.size() < 9)
                            throw std::runtime_error("triangle_normals entry malformed");
                        for (int k = 0; k < 9; ++k) hm->tri_normals[9 * i + k] = (float)t3.flat[k];
                    }
                }
            }
            // "triangle_uvs": [[u, v] x 3] per triangle (pybind_renderer.hpp:1295-1307)
            if (const Value* tu = mj.get("triangle_uvs")) {
                const int64_t n = tu->size();
                hm->tri_uvs.resize(6 * n);
                if (tu->kind == Value::NumArr) {
                    if (n > 0 && (tu->flat.size() != (size_t)(6 * n) || !tu->regular))
                        throw std::runtime_error("triangle_uvs must be 3 x [u, v] per triangle");
                    for (int64_t i = 0; i < 6 * n; ++i) hm->tri_uvs[i] = (float)tu->flat[i];
                } else {
                    for (int64_t i = 0; i < n; ++i) {
                        const Value& t3 = tu->arr[i];
                        if (t3.kind != Value::NumArr || t3.flat.size() < 6)
                            throw std::runtime_error("triangle_uvs entry malformed");
                        for (int k = 0; k < 6; ++k) hm->tri_uvs[6 * i + k] = (float)t3.flat[k];
                    }
                }
                d.num_tri_uvs = (int32_t)n;
            }
            bool smooth = false;
            if (const Value* sm = mj.get("smooth")) smooth = lucid::json::boolean(*sm);
            if (const Value* ts = mj.get("triangles")) {
                int64_t n = ts->size();
                hm->indices.resize(3 * n);
                if (ts->kind == Value::NumArr) {
                    if (n > 0 && (ts->shape.size() < 2 || ts->shape[1] < 3 || !ts->regular))
                        throw std::runtime_error("triangles must be index triples");
                    int64_t stride = n > 0 ? (int64_t)ts->flat.size() / n : 3;
                    for (int64_t i = 0; i < n; ++i)
                        for (int k = 0; k < 3; ++k) hm->indices[3 * i + k] = (int32_t)ts->flat[i * stride + k];
                } else {
                    for (int64_t i = 0; i < n; ++i)
                        for (int k = 0; k < 3; ++k) hm->indices[3 * i + k] = (int32_t)lucid::json::at(ts->arr[i], k);
                }
            }
            // Material defaults (pybind_renderer.hpp:1347-1350) and Mesh's
            // default-constructed Material (core/material.hpp:414-421) when absent.
            d.base_color[0] = d.base_color[1] = d.base_color[2] = 0.8f;
            d.metallic = 0.0f;
            d.roughness = 0.5f;
            d.transmission = 0.0f;
            d.ior = 1.45f;
            d.surface_linked = 1;
            if (const Value* mat = mj.get("material")) {
                if (const Value* bc = mat->get("base_color"))
                    for (int k = 0; k < 3; ++k) d.base_color[k] = f_at(*bc, k);
                if (const Value* v = mat->get("metallic")) d.metallic = f(*v);
                if (const Value* v = mat->get("roughness")) d.roughness = f(*v);
                if (const Value* em = mat->get("emission"))
                    for (int k = 0; k < 3; ++k) d.emission[k] = f_at(*em, k);
                if (const Value* v = mat->get("transmission")) d.transmission = f(*v);
                if (const Value* v = mat->get("ior")) d.ior = f(*v);
                d.surface_linked = surface_linked(*mat) ? 1 : 0;
                parse_node_tree(*mat, *hm, d);
                const Value* vol = mat->get("volume");
                if (vol && vol->is_object()) {
                    d.has_volume = 1;
                    d.vol_color[0] = d.vol_color[1] = d.vol_color[2] = 1.0f;
                    if (const Value* c = vol->get("color"); c && c->is_array() && c->size() >= 3)
                        for (int k = 0; k < 3; ++k) d.vol_color[k] = f_at(*c, k);
                    if (const Value* c = vol->get("absorption_color"); c && c->is_array() && c->size() >= 3)
                        for (int k = 0; k < 3; ++k) d.vol_absorption[k] = f_at(*c, k);
                    if (const Value* v = vol->get("density")) d.vol_density = f(*v);
                    if (const Value* v = vol->get("anisotropy")) d.vol_anisotropy = f(*v);
                    const Value* sg = vol->get("smoke_grid");
                    if (sg && sg->is_object()) {
                        const Value* dims = sg->get("dims");
                        const Value* wmin = sg->get("world_min");
                        const Value* wmax = sg->get("world_max");
                        const Value* path = sg->get("sidecar_path");
                        if (dims && dims->is_array() && dims->size() == 3 && wmin && wmin->is_array() && wmax &&
                            wmax->is_array() && path) {
                            int dd[3];
                            for (int k = 0; k < 3; ++k) {
                                dd[k] = (int)lucid::json::at(*dims, k);
                                d.grid_min[k] = f_at(*wmin, k);
                                d.grid_max[k] = f_at(*wmax, k);
                            }
                            size_t nvox = (size_t)dd[0] * dd[1] * dd[2];
                            std::ifstream in(lucid::json::string(*path), std::ios::binary);
                            if (in && dd[0] > 0 && dd[1] > 0 && dd[2] > 0) {
                                hm->grid.assign(nvox, 0.0f);
                                in.read(reinterpret_cast<char*>(hm->grid.data()), nvox * sizeof(float));
                                for (int k = 0; k < 3; ++k) d.grid_dims[k] = dd[k];
                            } else {
                                std::cerr << "[Volume] failed to open sidecar " << lucid::json::string(*path)
                                          << " -- heterogeneous grid skipped\n";
                            }
                        }
                    }
                }
            }
            d.smooth = smooth ? 1 : 0;
            d.num_vertices = (int32_t)(hm->vertices.size() / 3);
            d.vertices = hm->vertices.data();
            d.num_triangles = (int32_t)(hm->indices.size() / 3);
            d.indices = hm->indices.data();
            d.tri_normals = hm->tri_normals.empty() ? nullptr : hm->tri_normals.data();
            d.num_tri_normals = (int32_t)ntn;
            d.grid = hm->grid.empty() ? nullptr : hm->grid.data();
            d.tri_uvs = hm->tri_uvs.empty() ? nullptr : hm->tri_uvs.data();
            for (int64_t i = 0; i < 3LL * d.num_triangles; ++i)
                if (hm->indices[i] < 0 || hm->indices[i] >= d.num_vertices)
                    throw std::runtime_error("triangle index out of range");
            sc.meshes.push_back(std::move(hm));
        }
    } else if (meshes->size() != 0) {
        throw std::runtime_error("\"meshes\" entries must be objects");
    }
    const Value* lights = j.get("native_lights");
    if (!(lights && lights->kind == Value::Arr)) {
        lights = j.get("lights");
        if (!(lights && lights->kind == Value::Arr)) lights = nullptr;
    }
    if (lights) {
        for (const Value& lj : lights->arr) {
            lucid_light_desc l;
            std::memset(&l, 0, sizeof(l));
            std::string type = "POINT";
            if (const Value* t = lj.get("type")) type = lucid::json::string(*t);
            if (const Value* p = lj.get("position"))
                for (int k = 0; k < 3; ++k) l.position[k] = f_at(*p, k);
            if (const Value* dv = lj.get("direction")) {
                l.has_direction = 1;
                for (int k = 0; k < 3; ++k) l.direction[k] = f_at(*dv, k);
            }
            l.color[0] = l.color[1] = l.color[2] = 1.0f;
            if (const Value* c = lj.get("color"))
                for (int k = 0; k < 3; ++k) l.color[k] = f_at(*c, k);
            l.energy = 1.0f;
            if (const Value* e = lj.get("energy")) l.energy = f(*e);
            if (const Value* r = lj.get("radius")) l.radius = f(*r);
            l.spot_size = (float)(3.14159265358979323846 / 4.0f);
            if (const Value* s = lj.get("spot_size")) l.spot_size = f(*s);
            if (const Value* s = lj.get("spot_blend")) l.spot_blend = f(*s);
            l.size = 1.0f;
            if (const Value* s = lj.get("size")) l.size = f(*s);
            l.size_y = l.size;
            if (const Value* s = lj.get("size_y")) l.size_y = f(*s);
            std::string shape = "SQUARE";
            if (const Value* s = lj.get("shape")) shape = lucid::json::string(*s);
            l.shape = shape == "DISK" ? LUCID_SHAPE_DISK
                      : shape == "ELLIPSE" ? LUCID_SHAPE_ELLIPSE
                      : shape == "RECTANGLE" ? LUCID_SHAPE_RECTANGLE
                                             : LUCID_SHAPE_SQUARE;
            if (const Value* r = lj.get("right")) {
                l.has_right = 1;
                for (int k = 0; k < 3; ++k) l.right[k] = f_at(*r, k);
            }
            if (const Value* u = lj.get("up")) {
                l.has_up = 1;
                for (int k = 0; k < 3; ++k) l.up[k] = f_at(*u, k);
            }
            if (type == "POINT") l.type = LUCID_LIGHT_POINT;
            else if (type == "SUN") l.type = LUCID_LIGHT_SUN;
            else if (type == "SPOT") l.type = LUCID_LIGHT_SPOT;
            else if (type == "AREA") l.type = LUCID_LIGHT_AREA;
            else continue;  // the reference keeps a default Light; it is skipped here
            sc.lights.push_back(l);
        }
    }
    const Value* env = j.get("environment");
    if (env && env->is_object()) {
        if (const Value* c = env->get("color"); c && c->is_array())
            for (int k = 0; k < 3; ++k) sc.env_color[k] = f_at(*c, k);
        if (const Value* s = env->get("strength")) sc.env_strength = f(*s);
    }
}

int algo_code(const std::string& a) {  // pybind_renderer.hpp:301-327
    if (a == "simple") return LUCID_ALGO_SIMPLE;
    if (a == "mis") return LUCID_ALGO_MIS;
    if (a == "volume_simple") return LUCID_ALGO_VOLUME_SIMPLE;
    if (a == "volume_mis") return LUCID_ALGO_VOLUME_MIS;
    if (a == "guided" || a == "guided_mis") return LUCID_ALGO_GUIDED_MIS;  // extension
    return LUCID_ALGO_NEE;
}

// PyRenderer::TopGroupSortBy (pybind_renderer.hpp:1003-1006)
enum class TopGroupSortBy { Variance, Mean };

class Renderer {
public:
    Renderer() { std::cerr << "[PyRenderer] HIP wavefront path tracer (liblucid_hip)\n"; }
    ~Renderer() {
        if (async_) lucid_session_destroy(async_);
        if (scene_) lucid_scene_release(scene_);
        if (ctx_) lucid_destroy(ctx_);
    }

    void cancel() {
        cancel_.store(true);
        std::lock_guard<std::mutex> l(ctx_mu_);
        if (ctx_) lucid_cancel(ctx_);
    }
    bool is_cancelled() const { return cancel_.load(); }
    void reset_cancel() {
        cancel_.store(false);
        std::lock_guard<std::mutex> l(ctx_mu_);
        if (ctx_) lucid_reset_cancel(ctx_);
    }

    bool load_scene_json(const std::string& text) {
        auto sc = std::make_unique<HostScene>();
        try {
            parse_scene(text, *sc);
        } catch (const std::exception& e) {
            std::cerr << "[PyRenderer] Failed to load scene: " << e.what() << "\n";
            scene_loaded_ = false;
            return false;
        }
        std::lock_guard<std::mutex> l(render_mu_);
        host_ = std::move(sc);
        if (scene_) {
            lucid_scene_release(scene_);
            scene_ = nullptr;
        }
        scene_loaded_ = true;
        if (ctx_) lucid_guiding_reset(ctx_);  // a new scene needs a new guide
        std::cerr << "[PyRenderer] Scene loaded: " << host_->meshes.size() << " meshes\n";
        return true;
    }

    void set_camera(float px, float py, float pz, float dx, float dy, float dz, float ux, float uy, float uz,
                    float fov) {
        cam_.pos[0] = px; cam_.pos[1] = py; cam_.pos[2] = pz;
        cam_.dir[0] = dx; cam_.dir[1] = dy; cam_.dir[2] = dz;
        cam_.up[0] = ux; cam_.up[1] = uy; cam_.up[2] = uz;
        cam_.fov_deg = fov;
        camera_set_ = true;
    }
    void set_algorithm(const std::string& a) { algorithm_ = a; }
    // Extension: density-grid storage (32 = FP32, the reference's; 16 =
    // FP16); the scene is re-uploaded at the next render.
    void set_grid_precision(int bits) {
        if (bits != 16 && bits != 32) throw std::invalid_argument("grid precision must be 16 or 32 bits");
        std::lock_guard<std::mutex> l(render_mu_);
        if (bits == grid_bits_) return;
        grid_bits_ = bits;
        if (scene_) {
            lucid_scene_release(scene_);
            scene_ = nullptr;
        }
    }
    int get_grid_precision() const { return grid_bits_; }
    std::string get_algorithm() const { return algorithm_; }
    bool is_scene_loaded() const { return scene_loaded_; }
    bool is_camera_set() const { return camera_set_; }
    int get_mesh_count() const { return scene_loaded_ && host_ ? (int)host_->meshes.size() : 0; }

    // The loaded scene's material of mesh `mesh` at texture coordinates uv
    // ((n, 2)): getMaterialParams + getEmission (path_tracer.hpp:45-72)
    // through lucid_material_eval, host only.  (n, 10): albedo RGB,
    // emission RGB, metallic, roughness, transmission, ior.
    py::array_t<float> debug_material_eval(int mesh, py::array_t<float, py::array::c_style | py::array::forcecast> uv) {
        if (!scene_loaded_ || !host_ || mesh < 0 || mesh >= (int)host_->meshes.size())
            throw std::out_of_range("no such mesh");
        if (uv.ndim() != 2 || uv.shape(1) != 2) throw std::invalid_argument("uv must have shape (n, 2)");
        const int32_t n = (int32_t)uv.shape(0);
        py::array_t<float> out({(py::ssize_t)n, (py::ssize_t)10});
        char err[256];
        if (lucid_material_eval(&host_->meshes[mesh]->desc, n, uv.data(), out.mutable_data(), err, sizeof(err)) !=
            LUCID_OK)
            throw std::runtime_error(err);
        return out;
    }

    std::vector<float> render_tile(int tx, int ty, int tw, int th, int fw, int fh, int samples, int offset,
                                   int max_depth) {
        reset_cancel();
        if (tw < 0 || th < 0) throw std::length_error("negative tile size");
        std::vector<float> px((size_t)tw * th * 4, 0.0f);
        if (!scene_loaded_ || !camera_set_) {
            std::cerr << "[PyRenderer] Scene or camera not set\n";
            return px;
        }
        if (tw == 0 || th == 0) return px;
        std::lock_guard<std::mutex> l(render_mu_);
        ensure_scene();
        apply_diagnostics();
        lucid_render_params pr = params(tx, ty, tw, th, fw, fh, samples, offset, max_depth);
        int rc = lucid_render_tile(ctx_, scene_, &cam_, &pr, px.data());
        if (rc == LUCID_ERR_CANCELLED) {
            std::cerr << "[PyRenderer] Render cancelled\n";
            return px;
        }
        check(rc);
        return px;
    }

    // render_debug / render_debug_gpu (pybind_renderer.hpp:371-520): one
    // unjittered camera ray per pixel, coloured by `mode` ("normal",
    // "albedo", "emission", "volume"; anything else magenta), rows bottom
    // first.  Both run the same device pass here.
    std::vector<float> render_debug(int tx, int ty, int tw, int th, int fw, int fh, const std::string& mode) {
        reset_cancel();
        if (tw < 0 || th < 0) throw std::length_error("negative tile size");
        std::vector<float> px((size_t)tw * th * 4, 0.0f);
        if (!scene_loaded_ || !camera_set_ || tw == 0 || th == 0) return px;
        const int32_t m = mode == "normal" ? LUCID_DEBUG_NORMAL
                          : mode == "albedo" ? LUCID_DEBUG_ALBEDO
                          : mode == "emission" ? LUCID_DEBUG_EMISSION
                          : mode == "volume" ? LUCID_DEBUG_VOLUME
                                             : -1;
        std::lock_guard<std::mutex> l(render_mu_);
        ensure_scene();
        lucid_render_params pr = params(tx, ty, tw, th, fw, fh, 1, 0, 1);
        check(lucid_render_debug(ctx_, scene_, &cam_, &pr, m, px.data()));
        return px;
    }

    // -- async accumulator (pybind_renderer.hpp:689-855) --------------------
    // A session of the current scene, camera and algorithm accumulates on a
    // worker thread; poll returns its latest averaged snapshot.
    void render_start_async(int width, int height, int max_depth) {
        reset_cancel();
        std::lock_guard<std::mutex> la(async_mu_);
        stop_async_locked();
        if (!scene_loaded_ || !camera_set_ || width <= 0 || height <= 0) return;
        lucid_session* s = nullptr;
        {
            std::lock_guard<std::mutex> l(render_mu_);
            ensure_scene();
            check(lucid_session_create(ctx_, scene_, &cam_, width, height, std::max(1, max_depth),
                                       algo_code(algorithm_), &s));
        }
        const int rc = lucid_session_start_async(s);
        if (rc != LUCID_OK) {
            lucid_session_destroy(s);
            check(rc);
        }
        async_ = s;
        async_w_ = width;
        async_h_ = height;
    }
    // (samples, RGBA averaged, rows bottom first); samples 0 = no snapshot
    std::pair<int64_t, std::vector<float>> render_poll_async() {
        std::lock_guard<std::mutex> la(async_mu_);
        if (!async_) return {0, {}};
        int64_t n = 0;
        uint32_t rev = 0;
        check(lucid_session_poll_async(async_, nullptr, &n, &rev));
        if (n <= 0) return {0, {}};
        std::vector<float> px((size_t)async_w_ * async_h_ * 4);
        check(lucid_session_poll_async(async_, px.data(), &n, &rev));
        if (n <= 0) return {0, {}};
        return {n, std::move(px)};
    }
    void render_stop_async() {
        std::lock_guard<std::mutex> la(async_mu_);
        stop_async_locked();
    }
    void render_reset_async(int width, int height, int max_depth) {
        std::lock_guard<std::mutex> la(async_mu_);
        if (!async_ || !scene_loaded_ || !camera_set_) return;
        const int rc = lucid_session_reset_async(async_, &cam_, width, height, std::max(1, max_depth));
        if (rc != LUCID_OK)  // the reference logs a size mismatch and keeps running
            std::cerr << "[PyRenderer] reset_async failed: " << lucid_last_error(ctx_) << "\n";
    }
    bool is_render_async_running() {
        std::lock_guard<std::mutex> la(async_mu_);
        return async_ != nullptr;
    }
    int render_snapshot_revision_async() {
        std::lock_guard<std::mutex> la(async_mu_);
        if (!async_) return 0;
        int64_t n = 0;
        uint32_t rev = 0;
        check(lucid_session_poll_async(async_, nullptr, &n, &rev));
        return (int)rev;
    }

    // Extension: launch-shape autotuner on a full-frame render of the
    // current scene, camera and algorithm (the output is discarded).
    py::dict autotune(int width, int height, int samples, int max_depth) {
        require_ready();
        lucid_launch_shape b{};
        {
            py::gil_scoped_release rel;
            std::lock_guard<std::mutex> l(render_mu_);
            ensure_scene();
            apply_diagnostics();
            lucid_render_params pr = params(0, 0, width, height, width, height, samples, 0, max_depth);
            check(lucid_autotune(ctx_, scene_, &cam_, &pr, &b));
        }
        py::dict d;
        d["traverse_groups"] = b.traverse_groups;
        d["refill"] = b.refill;
        d["shade_groups"] = b.shade_groups;
        d["ray_order"] = b.ray_order;
        d["shade_order"] = b.shade_order;
        d["wave_paths"] = b.wave_paths;
        d["walk_min"] = b.walk_min;
        d["leaf_tris"] = b.leaf_tris;
        d["ms"] = b.ms;
        return d;
    }

    // Extension: per-sample radiance, shape (tile_h * tile_w, samples, 3).
    py::array_t<float> debug_render_samples(int tx, int ty, int tw, int th, int fw, int fh, int samples, int offset,
                                            int max_depth) {
        require_ready();
        std::vector<float> buf((size_t)tw * th * std::max(samples, 0) * 3);
        {
            py::gil_scoped_release rel;
            std::lock_guard<std::mutex> l(render_mu_);
            ensure_scene();
            lucid_render_params pr = params(tx, ty, tw, th, fw, fh, samples, offset, max_depth);
            check(lucid_render_samples(ctx_, scene_, &cam_, &pr, buf.data()));
        }
        py::array_t<float> a({(py::ssize_t)tw * th, (py::ssize_t)samples, (py::ssize_t)3});
        std::memcpy(a.mutable_data(), buf.data(), buf.size() * sizeof(float));
        return a;
    }

    // Extension: closest hit of every (jittered) camera ray; arrays of
    // tile_h * tile_w * samples entries, pixel-major, top row first.
    py::dict debug_primary_hits(int tx, int ty, int tw, int th, int fw, int fh, int samples, int offset) {
        require_ready();
        size_t n = (size_t)tw * th * std::max(samples, 0);
        py::array_t<int32_t> mesh(n), tri(n);
        py::array_t<float> u(n), v(n), t(n);
        int32_t *pm = mesh.mutable_data(), *pt = tri.mutable_data();
        float *pu = u.mutable_data(), *pv = v.mutable_data(), *ptt = t.mutable_data();
        {
            py::gil_scoped_release rel;
            std::lock_guard<std::mutex> l(render_mu_);
            ensure_scene();
            lucid_render_params pr = params(tx, ty, tw, th, fw, fh, samples, offset, 1);
            check(lucid_primary_hits(ctx_, scene_, &cam_, &pr, pm, pt, pu, pv, ptt));
        }
        py::dict d;
        d["mesh"] = mesh; d["tri"] = tri; d["u"] = u; d["v"] = v; d["t"] = t;
        return d;
    }

    py::dict last_render_stats() const {
        py::dict d;
        if (!ctx_) return d;
        lucid_stats s{};
        lucid_get_stats(ctx_, &s);
        d["rays_closest"] = s.rays_closest;
        d["rays_shadow"] = s.rays_shadow;
        d["paths"] = s.paths;
        d["waves"] = s.waves;
        d["stack_overflows"] = s.stack_overflows;
        d["ms_render"] = s.ms_render;
        return d;
    }
    void set_wave_paths(int64_t n) {
        std::lock_guard<std::mutex> l(render_mu_);
        ensure_ctx();
        check(lucid_set_wave_paths(ctx_, n));
    }

    // object names (pybind_renderer.hpp:1140-1253)
    void set_object_names(const std::vector<std::string>& n) { names_ = n; }
    std::string get_object_name(int32_t id) const {
        if (id < 0) return "Environment";
        if ((size_t)id < names_.size()) return names_[id];
        return "Object_" + std::to_string(id);
    }
    std::string object_path_string(const std::vector<int32_t>& ids) const {
        return object_path(ids, lucid::diag::kLightUnknown);
    }
    // object_path_string with the group's light type (pybind_renderer.hpp:1160-1253)
    std::string object_path(const std::vector<int32_t>& ids, uint8_t light_type) const {
        using namespace lucid::diag;
        if (ids.empty()) return "Camera";
        std::string path;
        size_t start = ids.size();
        int32_t last = ids[start - 1];
        if (light_type == kLightEnvironment) {
            path = "Environment";
            if (last == -1) start--;
        } else if (light_type == kLightEmissive) {
            path = get_object_name(last);
            start--;
        } else if (light_type == kLightPoint || light_type == kLightArea || light_type == kLightDirectional ||
                   light_type == kLightSpot) {
            path = light_type == kLightPoint ? "PointLight"
                   : light_type == kLightArea ? "AreaLight"
                   : light_type == kLightDirectional ? "SunLight"
                                                     : "SpotLight";
            if (last <= -2) start--;
        } else if (last <= -2) {
            path = "Light(?)";
            start--;
        } else if (last == -1) {
            path = "Environment(?)";
            start--;
        } else {
            path = "(terminated)";
        }
        for (size_t i = start; i > 0; --i) {
            path += " \xE2\x86\x92 ";
            path += get_object_name(ids[i - 1]);
        }
        path += " \xE2\x86\x92 Camera";
        return path;
    }

    // ---- learned path guiding (extension) ----
    void set_guiding(float alpha, int train_steps, int batch, float lr, int stride, uint32_t seed, int max_depth,
                     bool learn_selection) {
        std::lock_guard<std::mutex> l(render_mu_);
        ensure_ctx();
        lucid_guiding_config c{alpha, train_steps, batch, lr, stride, seed, max_depth, learn_selection ? 1 : 0};
        check(lucid_set_guiding(ctx_, &c));
    }
    void reset_guiding() {
        std::lock_guard<std::mutex> l(render_mu_);
        ensure_ctx();
        check(lucid_guiding_reset(ctx_));
    }
    py::dict guiding_stats() {
        py::dict d;
        std::lock_guard<std::mutex> l(render_mu_);
        ensure_ctx();
        lucid_guiding_stats st{};
        check(lucid_guiding_get_stats(ctx_, &st));
        d["samples_total"] = st.samples_total;
        d["train_steps"] = st.train_steps;
        d["last_loss"] = st.last_loss;
        d["active"] = (bool)st.active;
        d["num_params"] = st.num_params;
        lucid_stats rs{};
        lucid_get_stats(ctx_, &rs);
        d["ms_guide_infer"] = rs.ms_guide_infer;
        d["ms_guide_train"] = rs.ms_guide_train;
        d["ms_ray_binning"] = rs.ms_ray_binning;
        return d;
    }
    py::array_t<float> get_guiding_params() {
        std::lock_guard<std::mutex> l(render_mu_);
        ensure_ctx();
        lucid_guiding_stats st{};
        check(lucid_guiding_get_stats(ctx_, &st));
        py::array_t<float> a(st.num_params);
        check(lucid_guiding_get_params(ctx_, a.mutable_data(), st.num_params));
        return a;
    }
    void set_guiding_params(py::array_t<float, py::array::c_style | py::array::forcecast> a) {
        std::lock_guard<std::mutex> l(render_mu_);
        ensure_ctx();
        check(lucid_guiding_set_params(ctx_, a.data(), (int32_t)a.size()));
    }

    // ---- diagnostics (pybind_renderer.hpp:885-1135) ----
    struct PixelInfo {  // PyRenderer::PixelDiagnosticInfo
        bool valid = false;
        size_t sample_count = 0;
        float variance = 0.0f;
        size_t group_count = 0, outlier_count = 0;
        std::vector<float> mean_rgb{0.0f, 0.0f, 0.0f};
        std::vector<lucid::diag::GroupInfo> top_groups;
    };
    void enable_diagnostics(const lucid::diag::RecordingConfig& c) {
        std::lock_guard<std::mutex> l(render_mu_);
        diag_cfg_ = c;
        diag_on_ = true;
        drop_film();
    }
    void disable_diagnostics() {
        std::lock_guard<std::mutex> l(render_mu_);
        diag_on_ = false;
        drop_film();
    }
    bool is_diagnostics_enabled() const { return diag_on_; }
    void clear_diagnostics() {
        std::lock_guard<std::mutex> l(render_mu_);
        if (ctx_) check(lucid_diag_clear(ctx_));
    }
    lucid::diag::GlobalStats get_diagnostic_stats() {
        std::lock_guard<std::mutex> l(render_mu_);
        lucid::diag::FilmSource f;
        if (!film(f)) return lucid::diag::GlobalStats{};
        return lucid::diag::global_stats(f);
    }
    py::array_t<float> get_diagnostic_variance_map() {
        std::vector<float> map;
        {
            std::lock_guard<std::mutex> l(render_mu_);
            lucid::diag::FilmSource f;
            if (film(f)) {
                std::vector<float> var(f.sampled_w * f.sampled_h, 0.0f);
                f.for_each([&](size_t i, const lucid::diag::PixelData& p) {
                    var[i] = lucid::diag::total_variance(p.total);
                });
                map.assign(f.width * f.height, 0.0f);
                for (size_t y = 0; y < f.height; ++y)
                    for (size_t x = 0; x < f.width; ++x) {
                        long long k = f.index_of(x, y);
                        if (k >= 0) map[y * f.width + x] = var[(size_t)k];
                    }
            }
        }
        py::array_t<float> a((py::ssize_t)map.size());
        if (!map.empty()) std::memcpy(a.mutable_data(), map.data(), map.size() * sizeof(float));
        return a;
    }
    std::string export_diagnostic_json() {
        std::lock_guard<std::mutex> l(render_mu_);
        lucid::diag::FilmSource f;
        if (!film(f)) return "{}";
        return lucid::diag::metadata_json(f);
    }
    std::vector<lucid::diag::GroupInfo> get_top_variance_groups(size_t count) {
        std::lock_guard<std::mutex> l(render_mu_);
        lucid::diag::FilmSource f;
        if (!film(f)) return {};
        return lucid::diag::convert_refs(lucid::diag::top_groups(f, count, false));
    }
    // Extensions: the DiagnosticExporter queries (diagnostic_export.hpp:78-249)
    // over the device film, which the reference reaches only through a host
    // DiagnosticFilm.
    std::vector<lucid::diag::GroupInfo> get_top_mean_groups(size_t count) {
        std::lock_guard<std::mutex> l(render_mu_);
        lucid::diag::FilmSource f;
        if (!film(f)) return {};
        return lucid::diag::convert_refs(lucid::diag::top_groups(f, count, true));
    }
    std::vector<lucid::diag::CoarseInfo> get_diagnostic_coarse_stats() {
        std::lock_guard<std::mutex> l(render_mu_);
        lucid::diag::FilmSource f;
        if (!film(f)) return {};
        return lucid::diag::coarse_stats(f);
    }
    std::vector<lucid::diag::Suggestion> generate_diagnostic_suggestions() {
        std::lock_guard<std::mutex> l(render_mu_);
        lucid::diag::FilmSource f;
        if (!film(f)) return {};
        return lucid::diag::suggestions(f);
    }
    py::bytes export_diagnostic_binary() {
        std::string b;
        {
            std::lock_guard<std::mutex> l(render_mu_);
            lucid::diag::FilmSource f;
            if (film(f)) b = lucid::diag::export_binary(f);
        }
        return py::bytes(b);
    }
    PixelInfo get_pixel_diagnostic(size_t x, size_t y, const std::string& sort_by) {
        using namespace lucid::diag;
        PixelInfo info;
        std::lock_guard<std::mutex> l(render_mu_);
        FilmSource f;
        if (!film(f)) return info;
        const long long k = f.index_of(x, y);
        if (k < 0) return info;
        PixelData p;
        check(lucid_diag_read_pixels(ctx_, k, 1, &p));
        const size_t ng = (size_t)std::min<uint64_t>(p.num_groups, kMaxGroups);
        // sort_by_mean / sort_by_variance reorder the pixel's groups in place
        // (pixel_path_data.hpp:129-150)
        if (sort_by == "mean")
            std::sort(p.groups, p.groups + ng,
                      [](const Group& a, const Group& b) { return mean_luminance(a.stats) > mean_luminance(b.stats); });
        else
            std::sort(p.groups, p.groups + ng,
                      [](const Group& a, const Group& b) { return total_variance(a.stats) > total_variance(b.stats); });
        check(lucid_diag_write_pixels(ctx_, k, 1, &p));
        info.valid = true;
        info.sample_count = (size_t)p.total_samples;
        info.variance = total_variance(p.total);
        info.group_count = ng;
        info.outlier_count = (size_t)p.num_outliers;
        info.mean_rgb = {p.total.mean.x, p.total.mean.y, p.total.mean.z};
        for (size_t i = 0; i < ng; ++i) {
            const Group& g = p.groups[i];
            GroupInfo gi;
            gi.pixel_x = x; gi.pixel_y = y;
            gi.signature = signature(g);
            gi.signature_heckbert = signature_heckbert(g);
            for (int v = 0; v < g.depth; ++v) gi.object_ids.push_back(g.vertices[v].object_id);
            gi.sample_count = g.stats.count;
            gi.mean_luminance = mean_luminance(g.stats);
            gi.variance_luminance = total_variance(g.stats);
            gi.mean_rgb = {g.stats.mean.x, g.stats.mean.y, g.stats.mean.z};
            gi.variance_rgb = {g.stats.variance.x, g.stats.variance.y, g.stats.variance.z};
            gi.depth = g.depth;
            gi.coarse_type = g.coarse_type;
            gi.coarse_type_name = coarse_name(g.coarse_type);
            gi.strategy = g.strategy;
            gi.strategy_name = strategy_name(g.strategy);
            gi.object_path = object_path(gi.object_ids, g.light_type);
            gi.positions.push_back({cam_.pos[0], cam_.pos[1], cam_.pos[2]});
            for (int v = 0; v < g.depth; ++v) {
                gi.positions.push_back({g.positions[v].x, g.positions[v].y, g.positions[v].z});
                gi.normals.push_back({g.normals[v].x, g.normals[v].y, g.normals[v].z});
            }
            info.top_groups.push_back(gi);
        }
        return info;
    }
    // Extensions: the device film as raw PixelPathData<16,4> bytes
    // (sampled_h, sampled_w, 9576) and the FP64 per-strategy Welford
    // (sampled_h, sampled_w, 4, 7) = {n, mean rgb, M2 rgb} per strategy
    // (BSDF, NEE, MIS_BSDF, MIS_NEE).
    py::array_t<uint8_t> get_diagnostic_film_bytes() {
        std::vector<uint8_t> buf;
        size_t sw = 0, sh = 0;
        {
            std::lock_guard<std::mutex> l(render_mu_);
            lucid::diag::FilmSource f;
            if (film(f)) {
                sw = f.sampled_w; sh = f.sampled_h;
                buf.resize(sw * sh * sizeof(lucid::diag::PixelData));
                if (!buf.empty()) check(lucid_diag_read_pixels(ctx_, 0, (int64_t)(sw * sh), buf.data()));
            }
        }
        py::array_t<uint8_t> a({(py::ssize_t)sh, (py::ssize_t)sw, (py::ssize_t)sizeof(lucid::diag::PixelData)});
        if (!buf.empty()) std::memcpy(a.mutable_data(), buf.data(), buf.size());
        return a;
    }
    py::array_t<double> get_diagnostic_strategy_stats() {
        std::vector<double> buf;
        size_t sw = 0, sh = 0;
        {
            std::lock_guard<std::mutex> l(render_mu_);
            lucid::diag::FilmSource f;
            if (film(f)) {
                sw = f.sampled_w; sh = f.sampled_h;
                buf.resize(sw * sh * 28);
                if (!buf.empty()) check(lucid_diag_read_strategy(ctx_, 0, (int64_t)(sw * sh), buf.data()));
            }
        }
        py::array_t<double> a({(py::ssize_t)sh, (py::ssize_t)sw, (py::ssize_t)4, (py::ssize_t)7});
        if (!buf.empty()) std::memcpy(a.mutable_data(), buf.data(), buf.size() * sizeof(double));
        return a;
    }

    std::string repr() const {
        return "<lucidrenderer.Renderer scene_loaded=" + std::string(scene_loaded_ ? "True" : "False") +
               " meshes=" + std::to_string(get_mesh_count()) + " algorithm='" + algorithm_ + "'>";
    }

private:
    std::atomic<bool> cancel_{false};
    std::mutex ctx_mu_, render_mu_, async_mu_;  // lock order: async_mu_, render_mu_
    lucid_session* async_ = nullptr;
    int async_w_ = 0, async_h_ = 0;
    void stop_async_locked() {
        if (!async_) return;
        lucid_session_destroy(async_);  // stops the worker after its current wave
        async_ = nullptr;
    }
    lucid_ctx* ctx_ = nullptr;
    lucid_scene* scene_ = nullptr;
    std::unique_ptr<HostScene> host_;
    bool scene_loaded_ = false, camera_set_ = false;
    lucid_camera cam_{};
    int grid_bits_ = 32;
    std::string algorithm_ = "nee";
    std::vector<std::string> names_;
    bool diag_on_ = false;
    lucid::diag::RecordingConfig diag_cfg_;
    bool diag_applied_ = false;
    int diag_applied_sub_ = 0;

    // Device diagnostics state follows diag_on_ / diag_cfg_ at render time.
    void apply_diagnostics() {
        const int sub = diag_cfg_.subsample_factor;
        if (diag_on_ == diag_applied_ && (!diag_on_ || sub == diag_applied_sub_)) return;
        lucid_diag_config c{diag_on_ ? 1 : 0, sub};
        check(lucid_set_diagnostics(ctx_, &c));
        diag_applied_ = diag_on_;
        diag_applied_sub_ = sub;
    }
    void drop_film() {
        if (!ctx_) return;
        lucid_diag_config c{diag_on_ ? 1 : 0, diag_cfg_.subsample_factor ? diag_cfg_.subsample_factor : 1};
        check(lucid_set_diagnostics(ctx_, &c));
        diag_applied_ = diag_on_;
        diag_applied_sub_ = c.subsample;
    }
    // The device film as a FilmSource, or false when there is none.
    bool film(lucid::diag::FilmSource& f) {
        if (!ctx_) return false;
        lucid_diag_info inf{};
        check(lucid_diag_get_info(ctx_, &inf));
        if (!inf.valid) return false;
        f.width = (size_t)inf.width; f.height = (size_t)inf.height;
        f.sampled_w = (size_t)inf.sampled_width; f.sampled_h = (size_t)inf.sampled_height;
        f.total_recorded = (size_t)inf.total_recorded;
        f.config = diag_cfg_;
        f.config.subsample_factor = (uint8_t)inf.subsample;
        lucid_ctx* c = ctx_;
        f.visit = [c](size_t first, size_t count,
                      const std::function<void(size_t, const lucid::diag::PixelData&)>& fn) {
            std::vector<lucid::diag::PixelData> buf(count);
            int rc = lucid_diag_read_pixels(c, (int64_t)first, (int64_t)count, buf.data());
            if (rc != LUCID_OK) throw std::runtime_error(std::string("lucidrenderer: ") + lucid_last_error(c));
            for (size_t i = 0; i < count; ++i) fn(first + i, buf[i]);
        };
        return true;
    }

    void require_ready() {
        if (!scene_loaded_ || !camera_set_) throw std::runtime_error("scene or camera not set");
    }
    void ensure_ctx() {
        std::lock_guard<std::mutex> l(ctx_mu_);
        if (ctx_) return;
        lucid_ctx* c = nullptr;
        int rc = lucid_create(0, &c);
        if (rc != LUCID_OK)
            throw std::runtime_error("lucidrenderer: no usable HIP device (lucid_create returned " +
                                     std::to_string(rc) + "); this build has no CPU fallback");
        ctx_ = c;
    }
    void ensure_scene() {
        ensure_ctx();
        if (scene_) return;
        check(lucid_set_grid_format(ctx_, grid_bits_));
        std::vector<lucid_mesh_desc> md;
        for (auto& m : host_->meshes) md.push_back(m->desc);
        lucid_scene* s = nullptr;
        int rc = lucid_scene_create(ctx_, md.data(), (int32_t)md.size(), host_->lights.data(),
                                    (int32_t)host_->lights.size(), host_->env_color, host_->env_strength, &s);
        check(rc);
        scene_ = s;
    }
    lucid_render_params params(int tx, int ty, int tw, int th, int fw, int fh, int samples, int offset,
                               int max_depth) const {
        lucid_render_params p{};
        p.tile_x = tx; p.tile_y = ty; p.tile_w = tw; p.tile_h = th;
        p.full_w = fw; p.full_h = fh;
        p.samples = samples; p.sample_offset = offset; p.max_depth = max_depth;
        p.algorithm = algo_code(algorithm_);
        return p;
    }
    void check(int rc) {
        if (rc != LUCID_OK)
            throw std::runtime_error(std::string("lucidrenderer: ") + (ctx_ ? lucid_last_error(ctx_) : "error") +
                                     " (code " + std::to_string(rc) + ")");
    }
};

}  // namespace

PYBIND11_MODULE(lucidrenderer, m) {
    m.doc() = "Lucid Path Tracer - MI355X (HIP/gfx950) wavefront backend with the reference Python API";
    py::class_<Renderer>(m, "Renderer")
        .def(py::init<>(), "Create a new renderer instance")
        .def("cancel", &Renderer::cancel, "Request cancellation of the current render (any thread)")
        .def("is_cancelled", &Renderer::is_cancelled)
        .def("reset_cancel", &Renderer::reset_cancel)
        .def("load_scene_json", &Renderer::load_scene_json, py::arg("json_str"), "Load scene from JSON string")
        .def("set_camera", &Renderer::set_camera, py::arg("pos_x"), py::arg("pos_y"), py::arg("pos_z"),
             py::arg("dir_x"), py::arg("dir_y"), py::arg("dir_z"), py::arg("up_x"), py::arg("up_y"),
             py::arg("up_z"), py::arg("fov_deg"))
        .def("set_algorithm", &Renderer::set_algorithm, py::arg("algorithm"))
        .def("get_algorithm", &Renderer::get_algorithm)
        .def("autotune", &Renderer::autotune, py::arg("width"), py::arg("height"), py::arg("samples") = 4,
             py::arg("max_depth") = 8,
             "Extension: pick the fastest launch shape for renders like this one (results do not change)")
        .def("set_grid_precision", &Renderer::set_grid_precision, py::arg("bits"),
             "Extension: store density grids as FP32 (32, default) or FP16 (16)")
        .def("get_grid_precision", &Renderer::get_grid_precision)
        .def("set_object_names", &Renderer::set_object_names, py::arg("names"))
        .def("get_object_name", &Renderer::get_object_name, py::arg("object_id"))
        .def("object_path_string", &Renderer::object_path_string, py::arg("object_ids"))
        .def("render_tile", &Renderer::render_tile, py::call_guard<py::gil_scoped_release>(), py::arg("tile_x"),
             py::arg("tile_y"), py::arg("tile_w"), py::arg("tile_h"), py::arg("full_w"), py::arg("full_h"),
             py::arg("samples") = 1, py::arg("sample_offset") = 0, py::arg("max_depth") = 8,
             "Render a tile and return the RGBA float list (SUM over samples, bottom row first)")
        .def("render_start_async", &Renderer::render_start_async, py::call_guard<py::gil_scoped_release>(),
             py::arg("width"), py::arg("height"), py::arg("max_depth") = 8,
             "Start a background accumulation of the current scene and camera into an on-GPU sum "
             "(stops any prior session first)")
        .def(
            "render_poll_async",
            [](Renderer& self) {
                std::pair<int64_t, std::vector<float>> r;
                {
                    py::gil_scoped_release rel;
                    r = self.render_poll_async();
                }
                if (r.second.empty()) return py::make_tuple((int)r.first, py::array_t<float>{});
                py::array_t<float> a((py::ssize_t)r.second.size());
                std::memcpy(a.mutable_data(), r.second.data(), r.second.size() * sizeof(float));
                return py::make_tuple((int)r.first, std::move(a));
            },
            "Return (samples_completed, pixels_numpy): RGBA flat, bottom row first, alpha 1, averaged over "
            "`samples`; samples == 0 means no snapshot yet")
        .def("render_snapshot_revision_async", &Renderer::render_snapshot_revision_async,
             py::call_guard<py::gil_scoped_release>(), "Monotonic snapshot counter")
        .def("render_stop_async", &Renderer::render_stop_async, py::call_guard<py::gil_scoped_release>(),
             "Stop the background session. No-op if not running.")
        .def("render_reset_async", &Renderer::render_reset_async, py::call_guard<py::gil_scoped_release>(),
             py::arg("width"), py::arg("height"), py::arg("max_depth") = 8,
             "Non-blocking in-place reset of the running session with the current camera (same size only)")
        .def("is_render_async_running", &Renderer::is_render_async_running,
             py::call_guard<py::gil_scoped_release>(), "True iff a background session exists")
        .def("render_debug", &Renderer::render_debug, py::call_guard<py::gil_scoped_release>(), py::arg("tile_x"),
             py::arg("tile_y"), py::arg("tile_w"), py::arg("tile_h"), py::arg("full_w"), py::arg("full_h"),
             py::arg("mode"), "Render debug visualization: 'normal', 'albedo', or 'emission'")
        .def("render_debug_gpu", &Renderer::render_debug, py::call_guard<py::gil_scoped_release>(),
             py::arg("tile_x"), py::arg("tile_y"), py::arg("tile_w"), py::arg("tile_h"), py::arg("full_w"),
             py::arg("full_h"), py::arg("mode"), "GPU debug visualization (the same device pass as render_debug)")
        .def("render_tile_gpu", &Renderer::render_tile, py::call_guard<py::gil_scoped_release>(),
             py::arg("tile_x"), py::arg("tile_y"), py::arg("tile_w"), py::arg("tile_h"), py::arg("full_w"),
             py::arg("full_h"), py::arg("samples") = 1, py::arg("sample_offset") = 0, py::arg("max_depth") = 8,
             "Same as render_tile (every render runs on the GPU)")
        .def(
            "render_tile_numpy",
            [](Renderer& self, int tx, int ty, int tw, int th, int fw, int fh, int samples, int offset, int md) {
                std::vector<float> px;
                {
                    py::gil_scoped_release rel;
                    px = self.render_tile(tx, ty, tw, th, fw, fh, samples, offset, md);
                }
                py::array_t<float> a({(py::ssize_t)th, (py::ssize_t)tw, (py::ssize_t)4});
                std::memcpy(a.mutable_data(), px.data(), px.size() * sizeof(float));
                return a;
            },
            py::arg("tile_x"), py::arg("tile_y"), py::arg("tile_w"), py::arg("tile_h"), py::arg("full_w"),
            py::arg("full_h"), py::arg("samples") = 1, py::arg("sample_offset") = 0, py::arg("max_depth") = 8,
            "Render a tile and return a NumPy array of shape (height, width, 4)")
        .def("debug_render_samples", &Renderer::debug_render_samples, py::arg("tile_x"), py::arg("tile_y"),
             py::arg("tile_w"), py::arg("tile_h"), py::arg("full_w"), py::arg("full_h"), py::arg("samples") = 1,
             py::arg("sample_offset") = 0, py::arg("max_depth") = 8)
        .def("debug_primary_hits", &Renderer::debug_primary_hits, py::arg("tile_x"), py::arg("tile_y"),
             py::arg("tile_w"), py::arg("tile_h"), py::arg("full_w"), py::arg("full_h"), py::arg("samples") = 1,
             py::arg("sample_offset") = 0)
        .def("debug_material_eval", &Renderer::debug_material_eval, py::arg("mesh"), py::arg("uv"))
        .def("last_render_stats", &Renderer::last_render_stats)
        .def("set_wave_paths", &Renderer::set_wave_paths, py::arg("paths"))
        .def("is_scene_loaded", &Renderer::is_scene_loaded)
        .def("is_camera_set", &Renderer::is_camera_set)
        .def("get_mesh_count", &Renderer::get_mesh_count)
        // diagnostics (pybind_module.cpp:231-273)
        .def("enable_diagnostics", &Renderer::enable_diagnostics, py::arg("config"),
             "Enable diagnostic path recording with given configuration")
        .def("disable_diagnostics", &Renderer::disable_diagnostics, "Disable diagnostic recording")
        .def("is_diagnostics_enabled", &Renderer::is_diagnostics_enabled)
        .def("clear_diagnostics", &Renderer::clear_diagnostics, "Clear all recorded diagnostic data")
        .def("get_diagnostic_stats", &Renderer::get_diagnostic_stats, "Get global diagnostic statistics")
        .def("get_diagnostic_variance_map", &Renderer::get_diagnostic_variance_map,
             "Get per-pixel variance as 1D NumPy array (width * height)")
        .def("export_diagnostic_json", &Renderer::export_diagnostic_json, "Export diagnostic data as JSON string")
        .def("get_top_variance_groups", &Renderer::get_top_variance_groups, py::arg("count") = 10,
             "Get top N path groups by variance")
        .def("get_top_mean_groups", &Renderer::get_top_mean_groups, py::arg("count") = 10,
             "Extension: top N path groups of the device film by mean contribution")
        .def("get_diagnostic_coarse_stats", &Renderer::get_diagnostic_coarse_stats,
             "Extension: the 8 coarse path-type statistics of the device film")
        .def("generate_diagnostic_suggestions", &Renderer::generate_diagnostic_suggestions,
             "Extension: improvement suggestions from the device film")
        .def("export_diagnostic_binary", &Renderer::export_diagnostic_binary,
             "Extension: the device film in the PVAR v1 binary format")
        .def("get_pixel_diagnostic", &Renderer::get_pixel_diagnostic, py::arg("x"), py::arg("y"),
             py::arg("sort_by") = "variance",
             "Get diagnostic data for a specific pixel, sorted by 'variance' (default) or 'mean'")
        .def("set_guiding", &Renderer::set_guiding, py::arg("alpha") = 0.5f, py::arg("train_steps") = 16,
             py::arg("batch") = 32768, py::arg("learning_rate") = 1e-2f, py::arg("record_stride") = 32,
             py::arg("seed") = 1u, py::arg("max_depth") = 0, py::arg("learn_selection") = false,
             "Extension: configure learned path guiding (algorithm 'guided') and start from a fresh network")
        .def("reset_guiding", &Renderer::reset_guiding, "Extension: fresh guiding network")
        .def("guiding_stats", &Renderer::guiding_stats, "Extension: training state of the guiding network")
        .def("get_guiding_params", &Renderer::get_guiding_params)
        .def("set_guiding_params", &Renderer::set_guiding_params, py::arg("params"))
        .def("get_diagnostic_film_bytes", &Renderer::get_diagnostic_film_bytes,
             "Extension: the film as PixelPathData<16,4> bytes, shape (sampled_h, sampled_w, 9576)")
        .def("get_diagnostic_strategy_stats", &Renderer::get_diagnostic_strategy_stats,
             "Extension: FP64 Welford per film pixel and strategy, shape (sampled_h, sampled_w, 4, 7)")
        .def("__repr__", &Renderer::repr);

    using namespace lucid::diag;
    py::class_<RecordingConfig>(m, "PathRecordingConfig")
        .def(py::init<>(), "Create default recording configuration")
        .def_readwrite("subsample_factor", &RecordingConfig::subsample_factor)
        .def_readwrite("max_groups_per_pixel", &RecordingConfig::max_groups_per_pixel)
        .def_readwrite("max_outliers_per_pixel", &RecordingConfig::max_outliers_per_pixel)
        .def_readwrite("max_depth", &RecordingConfig::max_depth)
        .def_readwrite("variance_threshold", &RecordingConfig::variance_threshold)
        .def_readwrite("outlier_threshold", &RecordingConfig::outlier_threshold)
        .def_static("minimal", &RecordingConfig::minimal)
        .def_static("standard", &RecordingConfig::standard)
        .def_static("detailed", &RecordingConfig::detailed)
        .def("estimate_memory_bytes", &RecordingConfig::estimate_memory_bytes, py::arg("width"), py::arg("height"))
        .def("__repr__", [](const RecordingConfig& c) {
            return "<PathRecordingConfig subsample=" + std::to_string(c.subsample_factor) +
                   " max_groups=" + std::to_string(c.max_groups_per_pixel) + ">";
        });
    py::class_<GroupInfo>(m, "ExportedGroupInfo")
        .def(py::init<>())
        .def_readonly("pixel_x", &GroupInfo::pixel_x)
        .def_readonly("pixel_y", &GroupInfo::pixel_y)
        .def_readonly("signature", &GroupInfo::signature)
        .def_readonly("signature_heckbert", &GroupInfo::signature_heckbert)
        .def_readonly("object_ids", &GroupInfo::object_ids)
        .def_readonly("object_path", &GroupInfo::object_path)
        .def_readonly("sample_count", &GroupInfo::sample_count)
        .def_readonly("mean_luminance", &GroupInfo::mean_luminance)
        .def_readonly("variance_luminance", &GroupInfo::variance_luminance)
        .def_readonly("mean_rgb", &GroupInfo::mean_rgb)
        .def_readonly("variance_rgb", &GroupInfo::variance_rgb)
        .def_readonly("depth", &GroupInfo::depth)
        .def_readonly("coarse_type", &GroupInfo::coarse_type)
        .def_readonly("coarse_type_name", &GroupInfo::coarse_type_name)
        .def_readonly("strategy_name", &GroupInfo::strategy_name)
        .def_readonly("positions", &GroupInfo::positions)
        .def_readonly("normals", &GroupInfo::normals)
        .def("__repr__", [](const GroupInfo& g) {
            return "<ExportedGroupInfo signature='" + g.signature + "' heckbert='" + g.signature_heckbert + "'" +
                   " path='" + g.object_path + "'" + " strategy='" + g.strategy_name + "'" +
                   " mean=" + std::to_string(g.mean_luminance) + " var=" + std::to_string(g.variance_luminance) +
                   " positions=" + std::to_string(g.positions.size()) + ">";
        });
    py::enum_<TopGroupSortBy>(m, "TopGroupSortBy")
        .value("Variance", TopGroupSortBy::Variance, "Sort by variance (noisy paths first)")
        .value("Mean", TopGroupSortBy::Mean, "Sort by mean contribution (bright paths first)")
        .export_values();
    py::class_<Renderer::PixelInfo>(m, "PixelDiagnosticInfo")
        .def(py::init<>())
        .def_readonly("valid", &Renderer::PixelInfo::valid)
        .def_readonly("sample_count", &Renderer::PixelInfo::sample_count)
        .def_readonly("variance", &Renderer::PixelInfo::variance)
        .def_readonly("group_count", &Renderer::PixelInfo::group_count)
        .def_readonly("outlier_count", &Renderer::PixelInfo::outlier_count)
        .def_readonly("mean_rgb", &Renderer::PixelInfo::mean_rgb)
        .def_readonly("top_groups", &Renderer::PixelInfo::top_groups)
        .def("__repr__", [](const Renderer::PixelInfo& p) {
            if (!p.valid) return std::string("<PixelDiagnosticInfo invalid>");
            return "<PixelDiagnosticInfo samples=" + std::to_string(p.sample_count) +
                   " variance=" + std::to_string(p.variance) + " groups=" + std::to_string(p.group_count) + ">";
        });
    py::class_<Suggestion>(m, "DiagnosticSuggestion")
        .def(py::init<>())
        .def_readonly("category", &Suggestion::category)
        .def_readonly("severity", &Suggestion::severity)
        .def_readonly("message", &Suggestion::message)
        .def_readonly("action", &Suggestion::action)
        .def("__repr__", [](const Suggestion& s) {
            return "<DiagnosticSuggestion [" + s.severity + "] " + s.message + ">";
        });
    py::class_<GlobalStats>(m, "GlobalDiagnosticStats")
        .def(py::init<>())
        .def_readonly("total_samples", &GlobalStats::total_samples)
        .def_readonly("active_pixels", &GlobalStats::active_pixels)
        .def_readonly("total_groups", &GlobalStats::total_groups)
        .def_readonly("total_overflow", &GlobalStats::total_overflow)
        .def_readonly("total_variance", &GlobalStats::total_variance);
    py::class_<CoarseInfo>(m, "ExportedCoarseStats")
        .def(py::init<>())
        .def_readonly("type_id", &CoarseInfo::type_id)
        .def_readonly("name", &CoarseInfo::name)
        .def_readonly("sample_count", &CoarseInfo::sample_count)
        .def_readonly("mean_luminance", &CoarseInfo::mean_luminance)
        .def_readonly("variance_luminance", &CoarseInfo::variance_luminance)
        .def_readonly("mean_rgb", &CoarseInfo::mean_rgb, "Extension: the merged mean")
        .def_readonly("variance_rgb", &CoarseInfo::variance_rgb, "Extension: the merged (max) variance");
    py::class_<HostFilm>(m, "DiagnosticFilm")
        .def(py::init<size_t, size_t, const RecordingConfig&>(), py::arg("width"), py::arg("height"),
             py::arg("config"))
        .def("width", &HostFilm::width)
        .def("height", &HostFilm::height)
        .def("sampled_width", &HostFilm::sampled_width)
        .def("sampled_height", &HostFilm::sampled_height)
        .def("total_paths_recorded", &HostFilm::total_paths_recorded)
        .def("global_stats", [](const HostFilm& f) { return global_stats(f.source()); })
        .def("metadata_json", [](const HostFilm& f) { return metadata_json(f.source()); })
        .def("record_path", &HostFilm::record_path, py::arg("x"), py::arg("y"), py::arg("trace"))
        .def("clear", &HostFilm::clear)
        .def("pixel_bytes",
             [](const HostFilm& f, size_t x, size_t y) {
                 FilmSource src = f.source();
                 long long k = src.index_of(x, y);
                 if (k < 0) throw py::index_error("pixel outside the film");
                 std::string b;
                 src.visit((size_t)k, 1, [&](size_t, const PixelData& p) {
                     b.assign(reinterpret_cast<const char*>(&p), sizeof(p));
                 });
                 return py::bytes(b);
             },
             py::arg("x"), py::arg("y"), "Extension: PixelPathData<16,4> bytes of the film pixel holding (x, y)")
        .def("film_bytes",
             [](const HostFilm& f) {
                 FilmSource src = f.source();
                 std::string b;
                 b.reserve(src.sampled_w * src.sampled_h * sizeof(PixelData));
                 src.for_each([&](size_t, const PixelData& p) {
                     b.append(reinterpret_cast<const char*>(&p), sizeof(p));
                 });
                 return py::bytes(b);
             },
             "Extension: every film pixel as PixelPathData<16,4> bytes, row-major over the sampled grid");
    struct HostTracer {
        HostFilm film;
        bool enabled = true;
    };
    py::class_<HostTracer>(m, "DiagnosticPathTracer")
        .def(py::init([](size_t w, size_t h, const RecordingConfig& c) { return HostTracer{HostFilm(w, h, c), true}; }),
             py::arg("width"), py::arg("height"), py::arg("config"))
        .def("is_enabled", [](const HostTracer& t) { return t.enabled; })
        .def("set_enabled", [](HostTracer& t, bool e) { t.enabled = e; })
        .def("film", [](HostTracer& t) -> HostFilm& { return t.film; }, py::return_value_policy::reference_internal);
    struct Exporter {
        const HostFilm* film;
    };
    py::class_<Exporter>(m, "DiagnosticExporter")
        .def(py::init([](const HostFilm& f) { return Exporter{&f}; }), py::arg("film"), py::keep_alive<1, 2>())
        .def("get_global_stats", [](const Exporter& e) { return global_stats(e.film->source()); })
        .def("get_top_variance_groups",
             [](const Exporter& e, size_t n) { return convert_refs(top_groups(e.film->source(), n, false)); },
             py::arg("n") = 10)
        .def("get_top_mean_groups",
             [](const Exporter& e, size_t n) { return convert_refs(top_groups(e.film->source(), n, true)); },
             py::arg("n") = 10)
        .def("get_coarse_stats", [](const Exporter& e) { return coarse_stats(e.film->source()); })
        .def("generate_suggestions", [](const Exporter& e) { return suggestions(e.film->source()); })
        .def("export_metadata_json", [](const Exporter& e) { return metadata_json(e.film->source()); })
        .def("export_binary", [](const Exporter& e) { return py::bytes(export_binary(e.film->source())); });
    py::enum_<BsdfType>(m, "BsdfType")
        .value("Diffuse", kDiffuse)
        .value("Glossy", kGlossy)
        .value("Mirror", kMirror)
        .value("Glass", kGlass)
        .value("Emission", kEmission)
        .value("Environment", kEnvironment)
        .export_values();
    py::enum_<LightSource>(m, "LightSourceType")
        .value("Unknown", kLightUnknown)
        .value("Point", kLightPoint)
        .value("Area", kLightArea)
        .value("Directional", kLightDirectional)
        .value("Spot", kLightSpot)
        .value("Environment", kLightEnvironment)
        .value("Emissive", kLightEmissive)
        .export_values();
    py::class_<HostTrace>(m, "PathTrace")
        .def(py::init<>(), "Create empty path trace")
        .def_property_readonly("depth", [](const HostTrace& t) { return (size_t)t.n; })
        .def("get_contribution", [](const HostTrace& t) { return std::vector<float>{t.c.x, t.c.y, t.c.z}; });
    py::class_<HostRecorder>(m, "PathDiagnosticRecorder")
        .def(py::init<>())
        .def("begin_path", &HostRecorder::begin_path)
        .def("record_vertex",
             [](HostRecorder& r, int32_t obj, int16_t mat, BsdfType t, bool delta, bool ls) {
                 r.record_vertex(obj, mat, (uint8_t)t, delta, ls);
             },
             py::arg("object_id"), py::arg("material_id"), py::arg("bsdf_type"), py::arg("is_delta"),
             py::arg("is_light_sampled"))
        .def("record_environment_hit", &HostRecorder::record_environment_hit)
        .def("record_light_hit", [](HostRecorder& r, int32_t id, LightSource t) { r.record_light_hit(id, (uint8_t)t); },
             py::arg("light_id"), py::arg("light_source_type") = kLightUnknown)
        .def("record_emissive_hit", &HostRecorder::record_emissive_hit, py::arg("object_id"), py::arg("material_id"))
        .def("current_depth", &HostRecorder::current_depth)
        .def("end_path", &HostRecorder::end_path, py::arg("r"), py::arg("g"), py::arg("b"))
        .def("end_path_lum", [](HostRecorder& r, float l) { return r.end_path(l, l, l); }, py::arg("luminance"));

    m.attr("__version__") = "1.0.0";
    m.attr("openmp_enabled") = false;
    m.attr("dawn_enabled") = false;
    m.attr("hip_enabled") = true;
}
