// node_program.cpp -- compiles a material's shader-node tree into the
// postfix programs the shading kernel runs (GNodeOp, scene_pack.h).
//
// The value of every socket follows src/node_evaluator.cpp of the reference
// (LucidRenderer/cpp_renderer): evaluateNode :168-415, getAlbedoFromNodeTree
// :436-478, getEmissionFromNodeTree :501-565 and the scalar getters
// get{Transmission,IOR,Metallic,Roughness}FromNodeTree :579-740.  The
// reference walks the tree by name at every hit; here the walk happens once,
// at scene creation, and produces an expression over uv.  Subtrees that do
// not depend on uv are folded to constants with the reference's own float
// operations, in its order (this file is compiled with -ffp-contract=off),
// so a folded value is bit-identical to what the reference computes.  Only a
// Checker texture (:352-401) depends on uv.
//
// The reference recurses without bound (a link cycle never returns); here a
// chain of links deeper than kMaxLinkDepth is rejected with an error.

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "scene_pack.h"

namespace lucid {
namespace {

constexpr int kMaxLinkDepth = 64;

struct Rgb {
    float r, g, b;
};
const Rgb kGray = {0.8f, 0.8f, 0.8f};   // unknown node, missing socket
const Rgb kWhite = {1.0f, 1.0f, 1.0f};
const Rgb kBlack = {0.0f, 0.0f, 0.0f};

bool same(const char* a, const char* b) { return std::strcmp(a ? a : "", b) == 0; }
bool linked(const lucid_node_socket* s) { return s->linked_node != nullptr; }

// VEC4 reads its RGB part, VEC3 its xyz; any other kind is not a colour.
bool color_of(const lucid_node_socket* s, Rgb& c) {
    if (s->kind != LUCID_SOCK_VEC4 && s->kind != LUCID_SOCK_VEC3) return false;
    c = {s->value[0], s->value[1], s->value[2]};
    return true;
}

// One expression: a constant, or an op over child expressions.
struct Expr {
    int32_t op;   // NodeOpcode
    float k;      // CHECKER scale, MIX factor, SCALE factor
    Rgb c;        // NOP_CONST value
    int32_t x, y; // CHECKER color1 / color2, MIX A / B, SCALE operand
};

class TreeCompiler {
public:
    TreeCompiler(const lucid_node* nodes, int32_t n) : nodes_(nodes), n_(n) {}

    std::vector<Expr> ex;
    std::string err;

    // MaterialNode lookups (include/core/material.hpp:94-127): first match.
    const lucid_node* node(const char* name) const {
        for (int32_t i = 0; i < n_; ++i)
            if (same(nodes_[i].name, name ? name : "")) return &nodes_[i];
        return nullptr;
    }
    const lucid_node* output_node() const {
        for (int32_t i = 0; i < n_; ++i)
            if (same(nodes_[i].type, "ShaderNodeOutputMaterial")) return &nodes_[i];
        return nullptr;
    }
    static const lucid_node_socket* input(const lucid_node* nd, const char* name) {
        for (int32_t i = 0; i < nd->num_inputs; ++i)
            if (same(nd->inputs[i].name, name)) return &nd->inputs[i];
        return nullptr;
    }
    // The shader node the Material Output's Surface input is linked from,
    // with that link's socket name; null when there is none.
    const lucid_node_socket* surface_link() const {
        const lucid_node* out = output_node();
        if (!out) return nullptr;
        const lucid_node_socket* s = input(out, "Surface");
        return (s && linked(s)) ? s : nullptr;
    }

    int constant(Rgb c) {
        ex.push_back({NOP_CONST, 0.0f, c, -1, -1});
        return (int)ex.size() - 1;
    }
    bool is_const(int e) const { return ex[e].op == NOP_CONST; }

    // colorA * (1 - fac) + colorB * fac (node_evaluator.cpp:346)
    int mix(int a, int b, float fac) {
        if (a < 0 || b < 0) return -1;
        if (is_const(a) && is_const(b)) {
            const Rgb A = ex[a].c, B = ex[b].c;
            const float f1 = 1.0f - fac;
            return constant({A.r * f1 + B.r * fac, A.g * f1 + B.g * fac, A.b * f1 + B.b * fac});
        }
        ex.push_back({NOP_MIX, fac, kBlack, a, b});
        return (int)ex.size() - 1;
    }
    // color * strength (:261, :540, :549, :558)
    int scale(int c, float s) {
        if (c < 0) return -1;
        if (is_const(c)) {
            const Rgb C = ex[c].c;
            return constant({C.r * s, C.g * s, C.b * s});
        }
        ex.push_back({NOP_SCALE, s, kBlack, c, -1});
        return (int)ex.size() - 1;
    }
    int checker(int c1, int c2, float s) {
        if (c1 < 0 || c2 < 0) return -1;
        if (is_const(c1) && is_const(c2) && std::memcmp(&ex[c1].c, &ex[c2].c, sizeof(Rgb)) == 0) return c1;
        ex.push_back({NOP_CHECKER, s, kBlack, c1, c2});
        return (int)ex.size() - 1;
    }

    // A colour input: the linked node's value, the socket's colour, or
    // `fallback` when the socket holds no colour.
    int color_input(const lucid_node_socket* s, Rgb fallback, int depth) {
        if (!s) return constant(fallback);
        if (linked(s)) return eval(s->linked_node, s->linked_socket, depth + 1);
        Rgb c;
        return constant(color_of(s, c) ? c : fallback);
    }

    // evaluateNode (node_evaluator.cpp:168-415)
    int eval(const char* name, const char* socket, int depth) {
        if (depth > kMaxLinkDepth) {
            err = "node tree: a chain of links deeper than 64 (or a cycle) below \"" +
                  std::string(name ? name : "") + "\"";
            return -1;
        }
        const lucid_node* nd = node(name);
        if (!nd) return constant(kGray);
        const char* type = nd->type;
        if (same(type, "ShaderNodeBsdfPrincipled")) {
            if (same(socket, "BSDF")) {
                const lucid_node_socket* s = input(nd, "Base Color");
                if (!s) return constant(kGray);
                return color_input(s, kGray, depth);
            }
            if (same(socket, "Base Color")) {
                const lucid_node_socket* s = input(nd, "Base Color");
                Rgb c;
                if (s && !linked(s) && color_of(s, c)) return constant(c);
                return constant(kGray);
            }
        } else if (same(type, "ShaderNodeEmission")) {
            if (same(socket, "Emission")) {
                const lucid_node_socket* cs = input(nd, "Color");
                const lucid_node_socket* ss = input(nd, "Strength");
                int color = color_input(cs, kWhite, depth);
                float strength = 1.0f;
                if (ss && !linked(ss) && ss->kind == LUCID_SOCK_FLOAT) strength = ss->value[0];
                return scale(color, strength);
            }
        } else if (same(type, "ShaderNodeBsdfDiffuse")) {
            if (same(socket, "BSDF")) {
                const lucid_node_socket* s = input(nd, "Color");
                if (!s) return constant(kGray);
                return color_input(s, kGray, depth);
            }
        } else if (same(type, "ShaderNodeRGB")) {
            for (int32_t i = 0; i < nd->num_outputs; ++i) {
                const lucid_node_socket& o = nd->outputs[i];
                Rgb c;
                if (same(o.name, "Color") && color_of(&o, c)) return constant(c);
            }
            return constant(kWhite);
        } else if (same(type, "ShaderNodeMix") || same(type, "ShaderNodeMixRGB")) {
            const lucid_node_socket* fs = input(nd, "Fac");
            float fac = 0.5f;
            if (fs && !linked(fs) && fs->kind == LUCID_SOCK_FLOAT) fac = fs->value[0];
            int a = color_input(input(nd, "A"), kBlack, depth);
            int b = color_input(input(nd, "B"), kWhite, depth);
            return mix(a, b, fac);
        } else if (same(type, "ShaderNodeTexChecker")) {
            int c1 = color_input(input(nd, "Color1"), kGray, depth);
            int c2 = color_input(input(nd, "Color2"), {0.2f, 0.2f, 0.2f}, depth);
            const lucid_node_socket* ss = input(nd, "Scale");
            float s = 5.0f;
            if (ss && !linked(ss) && ss->kind == LUCID_SOCK_FLOAT) s = ss->value[0];
            return checker(c1, c2, s);
        } else if (same(type, "ShaderNodeOutputMaterial")) {
            const lucid_node_socket* s = input(nd, "Surface");
            if (s && linked(s)) return eval(s->linked_node, s->linked_socket, depth + 1);
        }
        return constant(kGray);
    }

    // getAlbedoFromNodeTree (:436-478)
    int albedo() {
        const lucid_node_socket* s = surface_link();
        if (!s) return constant(kGray);
        return eval(s->linked_node, s->linked_socket, 1);
    }

    // getEmissionFromNodeTree (:501-565)
    int emission() {
        const lucid_node_socket* s = surface_link();
        if (!s) return constant(kBlack);
        const lucid_node* sh = node(s->linked_node);
        if (sh && same(sh->type, "ShaderNodeEmission")) return eval(s->linked_node, "Emission", 1);
        if (sh && same(sh->type, "ShaderNodeBsdfPrincipled")) {
            const lucid_node_socket* es = input(sh, "Emission Color");
            if (!es) es = input(sh, "Emission");
            if (es) {
                // the strength is read whether or not its socket is linked
                const lucid_node_socket* st = input(sh, "Emission Strength");
                const float strength = (st && st->kind == LUCID_SOCK_FLOAT) ? st->value[0] : 0.0f;
                if (linked(es)) return scale(eval(es->linked_node, es->linked_socket, 1), strength);
                Rgb c;
                if (color_of(es, c)) return scale(constant(c), strength);
            }
        }
        return constant(kBlack);
    }

    // The Principled BSDF's unlinked FLOAT input `name` (or `alt`), else `dflt`
    // (get{Transmission,IOR,Metallic,Roughness}FromNodeTree).
    float principled_scalar(const char* name, const char* alt, float dflt) const {
        const lucid_node_socket* s = surface_link();
        if (!s) return dflt;
        const lucid_node* sh = node(s->linked_node);
        if (!sh || !same(sh->type, "ShaderNodeBsdfPrincipled")) return dflt;
        const lucid_node_socket* in = input(sh, name);
        if (!in && alt) in = input(sh, alt);
        if (in && !linked(in) && in->kind == LUCID_SOCK_FLOAT) return in->value[0];
        return dflt;
    }

    // Evaluation slots an expression needs (Sethi-Ullman numbering).
    int slots(int e) const {
        const Expr& x = ex[e];
        if (x.op == NOP_CONST) return 1;
        if (x.op == NOP_SCALE) return slots(x.x);
        int a = slots(x.x), b = slots(x.y);
        return a == b ? a + 1 : (a > b ? a : b);
    }
    // Postfix code; a binary op evaluates its hungrier operand first.
    void emit(int e, std::vector<GNodeOp>& out) const {
        const Expr& x = ex[e];
        if (x.op == NOP_CONST) {
            out.push_back({NOP_CONST, x.c.r, x.c.g, x.c.b});
        } else if (x.op == NOP_SCALE) {
            emit(x.x, out);
            out.push_back({NOP_SCALE, x.k, 0.0f, 0.0f});
        } else {
            const bool swap = slots(x.y) > slots(x.x);
            emit(swap ? x.y : x.x, out);
            emit(swap ? x.x : x.y, out);
            out.push_back({x.op | (swap ? kNopSwapped : 0), x.k, 0.0f, 0.0f});
        }
    }

private:
    const lucid_node* nodes_;
    int32_t n_;
};

// A program for expression e, or an empty one with the constant in rgb.
bool lower(const TreeCompiler& tc, int e, float rgb[3], std::vector<GNodeOp>& prog, std::string& err) {
    prog.clear();
    if (tc.is_const(e)) {
        rgb[0] = tc.ex[e].c.r; rgb[1] = tc.ex[e].c.g; rgb[2] = tc.ex[e].c.b;
        return true;
    }
    if (tc.slots(e) > kNodeStack) {
        err = "node tree: the texture expression needs more than " + std::to_string(kNodeStack) +
              " evaluation slots";
        return false;
    }
    tc.emit(e, prog);
    prog.push_back({NOP_END, 0.0f, 0.0f, 0.0f});
    rgb[0] = rgb[1] = rgb[2] = 0.0f;
    return true;
}

}  // namespace

bool compile_material(const lucid_mesh_desc& d, CompiledMaterial& out, std::string& err) {
    out.albedo_prog.clear();
    out.emission_prog.clear();
    for (int k = 0; k < 3; ++k) {
        out.albedo[k] = d.base_color[k];
        out.emission[k] = d.emission[k];
    }
    out.metallic = d.metallic;
    out.roughness = d.roughness;
    out.transmission = d.transmission;
    out.ior = d.ior;
    // useNodes = "use_nodes" && the tree has a node (pybind_renderer.hpp:1381-1384)
    if (!d.use_nodes || d.num_nodes <= 0) return true;
    if (!d.nodes) {
        err = "use_nodes with num_nodes > 0 needs a node array";
        return false;
    }
    for (int32_t i = 0; i < d.num_nodes; ++i) {
        const lucid_node& n = d.nodes[i];
        if ((n.num_inputs > 0 && !n.inputs) || (n.num_outputs > 0 && !n.outputs) || n.num_inputs < 0 ||
            n.num_outputs < 0) {
            err = "node " + std::to_string(i) + ": socket count without a socket array";
            return false;
        }
    }
    TreeCompiler tc(d.nodes, d.num_nodes);
    const int a = tc.albedo();
    const int e = a < 0 ? -1 : tc.emission();
    if (a < 0 || e < 0) {
        err = tc.err;
        return false;
    }
    if (!lower(tc, a, out.albedo, out.albedo_prog, err)) return false;
    if (!lower(tc, e, out.emission, out.emission_prog, err)) return false;
    out.metallic = tc.principled_scalar("Metallic", nullptr, 0.0f);
    out.roughness = tc.principled_scalar("Roughness", nullptr, 0.5f);
    out.transmission = tc.principled_scalar("Transmission", "Transmission Weight", 0.0f);
    out.ior = tc.principled_scalar("IOR", nullptr, 1.45f);
    return true;
}

// Checker cell parity of one coordinate: floor(x) mod 2 with the
// reference's negative fix-up, i.e. the low bit of floor(x).  Floats of
// magnitude 2^24 and above are even integers; NaN reads as even, as the
// x86 conversion of NaN to INT_MIN makes it in the reference.
static inline int checker_parity(float x) {
    const float f = std::floor(x);
    return (std::fabs(f) < 16777216.0f) ? ((int)f & 1) : 0;
}

void run_node_program(const GNodeOp* ops, float u, float v, float rgb[3]) {
    float st[kNodeStack][3];
    int sp = 0;
    for (const GNodeOp* o = ops; (o->op & 0xff) != NOP_END; ++o) {
        const int code = o->op & 0xff;
        if (code == NOP_CONST) {
            st[sp][0] = o->a; st[sp][1] = o->b; st[sp][2] = o->c;
            ++sp;
        } else if (code == NOP_SCALE) {
            for (int k = 0; k < 3; ++k) st[sp - 1][k] = st[sp - 1][k] * o->a;
        } else {
            const bool sw = (o->op & kNopSwapped) != 0;
            const float* x = sw ? st[sp - 1] : st[sp - 2];  // color1 / A
            const float* y = sw ? st[sp - 2] : st[sp - 1];  // color2 / B
            float r[3];
            if (code == NOP_CHECKER) {
                const bool first = checker_parity(u * o->a) == checker_parity(v * o->a);
                for (int k = 0; k < 3; ++k) r[k] = first ? x[k] : y[k];
            } else {
                const float f1 = 1.0f - o->a;
                for (int k = 0; k < 3; ++k) r[k] = x[k] * f1 + y[k] * o->a;
            }
            --sp;
            for (int k = 0; k < 3; ++k) st[sp - 1][k] = r[k];
        }
    }
    for (int k = 0; k < 3; ++k) rgb[k] = st[0][k];
}

}  // namespace lucid

// ------------------------------------------------------------------ C ABI --
extern "C" int lucid_material_eval(const lucid_mesh_desc* mesh, int32_t count, const float* uv, float* out,
                                   char* err, int32_t err_len) {
    auto fail = [&](const std::string& msg) {
        if (err && err_len > 0) {
            std::strncpy(err, msg.c_str(), (size_t)err_len - 1);
            err[err_len - 1] = '\0';
        }
        return LUCID_ERR_INVALID;
    };
    if (err && err_len > 0) err[0] = '\0';
    if (!mesh || count < 0 || (count > 0 && (!uv || !out))) return fail("null argument");
    lucid::CompiledMaterial cm;
    std::string msg;
    if (!lucid::compile_material(*mesh, cm, msg)) return fail(msg);
    for (int32_t i = 0; i < count; ++i) {
        float* o = out + 10 * (size_t)i;
        const float u = uv[2 * (size_t)i], v = uv[2 * (size_t)i + 1];
        if (cm.albedo_prog.empty()) {
            for (int k = 0; k < 3; ++k) o[k] = cm.albedo[k];
        } else {
            lucid::run_node_program(cm.albedo_prog.data(), u, v, o);
        }
        if (cm.emission_prog.empty()) {
            for (int k = 0; k < 3; ++k) o[3 + k] = cm.emission[k];
        } else {
            lucid::run_node_program(cm.emission_prog.data(), u, v, o + 3);
        }
        o[6] = cm.metallic;
        o[7] = (cm.roughness < 0.01f) ? 0.01f : cm.roughness;  // std::max(roughness, MIN_ROUGHNESS)
        o[8] = cm.transmission;
        o[9] = cm.ior;
    }
    return LUCID_OK;
}
