// Host-side glTF 2.0 binary (.glb) ingest: the arrays loadGltfScene
// (src/lib/components/simulationViewer/loadScene.ts:31-251) hands to the collider and scatter
// managers, computed in double precision and stored as FP32.
//   - meshes are visited depth-first in scene order (loadScene.ts:19-25 traverseChildren); a glTF
//     mesh with several primitives is one three.js Mesh per primitive, in primitive order, before
//     the node's child nodes;
//   - world matrix = parent world x local; local = T R S of the node (or its matrix);
//   - a position p becomes (x, z, y) of matrixWorld p (loadScene.ts:27-30, the y/z swap);
//   - normals: normal matrix (inverse transpose of the upper 3x3), normalised, NOT swapped (as
//     loadScene.ts:182-187 pushes them);
//   - material index of a vertex: the first-appearance index of its material's base-colour
//     texture, 0 without one (loadScene.ts:150-158);
//   - indices offset by the vertices of the previous meshes; one collider object (AABB, first index,
//     index count) per mesh; the scatter mesh is the vertex of every index in order
//     (loadScene.ts:166-179).
// Supported: embedded BIN chunk buffers, float positions / normals, float or normalised u8 / u16
// texture coordinates (normalised ones copied as their raw integers, as loadScene.ts:190-193 copies
// the attribute array), u8 / u16 / u32 indices, node matrices and TRS. Not supported (an error):
// external or data-URI buffers, sparse accessors, non-indexed primitives (loadScene.ts reads
// geometry.index unconditionally).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace mpm_gltf {

// ------------------------------------------------------------------ minimal JSON
struct JVal {
    enum Type { Null, Bool, Num, Str, Arr, Obj } t = Null;
    double num = 0.0;
    bool b = false;
    std::string s;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const char* k) const {
        if (t != Obj) return nullptr;
        for (const auto& kv : obj) if (kv.first == k) return &kv.second;
        return nullptr;
    }
    const JVal* at(size_t i) const { return (t == Arr && i < arr.size()) ? &arr[i] : nullptr; }
    double number(const char* k, double dflt) const { const JVal* v = get(k); return (v && v->t == Num) ? v->num : dflt; }
};

// A JSON number that is a non-negative integer below `limit`: glTF indices, counts, offsets and
// strides. Rejects fractions, negatives, NaN and values beyond the limit before any cast.
inline bool as_uint(const JVal* v, double limit, size_t& out) {
    if (!v || v->t != JVal::Num) return false;
    const double d = v->num;
    if (!(d >= 0.0) || !(d < limit) || d != std::floor(d)) return false;
    out = (size_t)d;
    return true;
}
// An optional such field of an object: absent -> dflt.
inline bool opt_uint(const JVal* obj, const char* key, size_t dflt, double limit, size_t& out) {
    const JVal* v = obj ? obj->get(key) : nullptr;
    if (!v) { out = dflt; return true; }
    return as_uint(v, limit, out);
}
inline size_t arr_len(const JVal* v) { return (v && v->t == JVal::Arr) ? v->arr.size() : 0; }

struct JParser {
    const char* p;
    const char* e;
    std::string err;
    void ws() { while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p; }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if ((size_t)(e - p) < n || std::memcmp(p, w, n) != 0) return false;
        p += n;
        return true;
    }
    bool str(std::string& out) {
        if (p >= e || *p != '"') return false;
        ++p;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                if (++p >= e) return false;
                const char c = *p++;
                switch (c) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (e - p < 4) return false;
                        unsigned cp = (unsigned)std::strtoul(std::string(p, 4).c_str(), nullptr, 16);
                        p += 4;
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                        else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
                        break;
                    }
                    default: out += c;
                }
            } else {
                out += *p++;
            }
        }
        if (p >= e) return false;
        ++p;
        return true;
    }
    bool value(JVal& v, int depth) {
        if (depth > 64) { err = "JSON nested too deeply"; return false; }
        ws();
        if (p >= e) return false;
        const char c = *p;
        if (c == '{') {
            v.t = JVal::Obj; ++p; ws();
            if (p < e && *p == '}') { ++p; return true; }
            for (;;) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p >= e || *p != ':') return false;
                ++p;
                JVal x;
                if (!value(x, depth + 1)) return false;
                v.obj.emplace_back(std::move(k), std::move(x));
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == '}') { ++p; return true; }
                return false;
            }
        }
        if (c == '[') {
            v.t = JVal::Arr; ++p; ws();
            if (p < e && *p == ']') { ++p; return true; }
            for (;;) {
                JVal x;
                if (!value(x, depth + 1)) return false;
                v.arr.push_back(std::move(x));
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == ']') { ++p; return true; }
                return false;
            }
        }
        if (c == '"') { v.t = JVal::Str; return str(v.s); }
        if (lit("true")) { v.t = JVal::Bool; v.b = true; return true; }
        if (lit("false")) { v.t = JVal::Bool; v.b = false; return true; }
        if (lit("null")) { v.t = JVal::Null; return true; }
        std::string num;
        while (p < e && (std::strchr("+-0123456789.eE", *p) != nullptr)) num += *p++;
        if (num.empty()) return false;
        char* end = nullptr;
        v.t = JVal::Num;
        v.num = std::strtod(num.c_str(), &end);
        return end && *end == '\0';
    }
};

// ------------------------------------------------------------------ matrices (column-major, like three.js)
struct M4 { double e[16]; };
inline M4 ident4() { M4 m{}; m.e[0] = m.e[5] = m.e[10] = m.e[15] = 1.0; return m; }
// a * b (multiplyMatrices: sums of four products, left to right)
inline M4 mul4(const M4& a, const M4& b) {
    M4 r;
    for (int c = 0; c < 4; ++c)
        for (int rr = 0; rr < 4; ++rr)
            r.e[4 * c + rr] = a.e[rr] * b.e[4 * c] + a.e[4 + rr] * b.e[4 * c + 1] + a.e[8 + rr] * b.e[4 * c + 2] + a.e[12 + rr] * b.e[4 * c + 3];
    return r;
}
// T R S from translation, unit quaternion (x, y, z, w) and scale (Matrix4.compose form)
inline M4 compose(const double t[3], const double q[4], const double s[3]) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    const double x2 = x + x, y2 = y + y, z2 = z + z;
    const double xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
    const double wx = w * x2, wy = w * y2, wz = w * z2;
    M4 m;
    m.e[0] = (1 - (yy + zz)) * s[0]; m.e[1] = (xy + wz) * s[0]; m.e[2] = (xz - wy) * s[0]; m.e[3] = 0;
    m.e[4] = (xy - wz) * s[1]; m.e[5] = (1 - (xx + zz)) * s[1]; m.e[6] = (yz + wx) * s[1]; m.e[7] = 0;
    m.e[8] = (xz + wy) * s[2]; m.e[9] = (yz - wx) * s[2]; m.e[10] = (1 - (xx + yy)) * s[2]; m.e[11] = 0;
    m.e[12] = t[0]; m.e[13] = t[1]; m.e[14] = t[2]; m.e[15] = 1;
    return m;
}
// applyMatrix4 followed by the (x, z, y) swap of loadScene.ts:27-30
inline void xform_swap(const M4& m, const float* p, double out[3]) {
    const double x = p[0], y = p[1], z = p[2];
    const double w = 1.0 / (m.e[3] * x + m.e[7] * y + m.e[11] * z + m.e[15]);
    const double X = (m.e[0] * x + m.e[4] * y + m.e[8] * z + m.e[12]) * w;
    const double Y = (m.e[1] * x + m.e[5] * y + m.e[9] * z + m.e[13]) * w;
    const double Z = (m.e[2] * x + m.e[6] * y + m.e[10] * z + m.e[14]) * w;
    out[0] = X; out[1] = Z; out[2] = Y;
}
// normal matrix: inverse transpose of the upper 3x3 (row-major n[3][3])
inline bool normal_matrix(const M4& m, double n[3][3]) {
    const double a = m.e[0], b = m.e[4], c = m.e[8], d = m.e[1], e = m.e[5], f = m.e[9], g = m.e[2], h = m.e[6], i = m.e[10];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    if (det == 0.0) return false;
    const double inv = 1.0 / det;
    // inverse = adj / det; its transpose is cof / det
    n[0][0] = A * inv; n[0][1] = B * inv; n[0][2] = C * inv;
    n[1][0] = -(b * i - c * h) * inv; n[1][1] = (a * i - c * g) * inv; n[1][2] = -(a * h - b * g) * inv;
    n[2][0] = (b * f - c * e) * inv; n[2][1] = -(a * f - c * d) * inv; n[2][2] = (a * e - b * d) * inv;
    return true;
}

// ------------------------------------------------------------------ loader
struct Scene {
    std::vector<float> positions, normals, uvs, soup;
    std::vector<uint32_t> material_indices, indices;
    struct Object { float mn[3], mx[3]; uint32_t start, count; };
    std::vector<Object> objects;
};

struct Loader {
    const uint8_t* bin = nullptr;
    size_t bin_len = 0;
    JVal doc;
    std::string err;
    std::vector<int> tex_of_texture_index;   // glTF texture index -> first-appearance order
    int n_textures = 0;

    bool fail(const std::string& m) { if (err.empty()) err = m; return false; }

    // accessor elements as doubles; normalised integer attributes keep their raw integer values
    // (loadScene.ts:190-193 copies the attribute array as it is). Every count, offset and stride
    // is checked as an integer before use, and the byte range of the last element is checked
    // against the buffer view without an expression that can wrap.
    bool accessor(size_t idx, const char* want_type, std::vector<double>& out, int& ncomp, bool allow_norm_int) {
        const JVal* accs = doc.get("accessors");
        const JVal* a = accs ? accs->at(idx) : nullptr;
        if (!a) return fail("accessor index out of range");
        if (a->get("sparse")) return fail("sparse accessors are not supported");
        const JVal* ty = a->get("type");
        const std::string type = ty && ty->t == JVal::Str ? ty->s : "";
        ncomp = type == "SCALAR" ? 1 : type == "VEC2" ? 2 : type == "VEC3" ? 3 : type == "VEC4" ? 4 : 0;
        if (!ncomp || (want_type && type != want_type)) return fail("unexpected accessor type " + type);
        size_t ct = 0, count = 0, acc_off = 0;
        const double lim = (double)bin_len + 1.0;
        if (!opt_uint(a, "componentType", 0, 65536.0, ct)) return fail("invalid componentType");
        if (!opt_uint(a, "count", 0, lim, count)) return fail("invalid accessor count");
        if (!opt_uint(a, "byteOffset", 0, lim, acc_off)) return fail("invalid accessor byteOffset");
        const size_t csize = ct == 5126 || ct == 5125 ? 4 : ct == 5123 || ct == 5122 ? 2 : ct == 5121 || ct == 5120 ? 1 : 0;
        if (!csize) return fail("unsupported component type");
        if (ct != 5126 && ct != 5125 && ct != 5123 && ct != 5121) return fail("signed integer accessors are not supported");
        const JVal* bvs = doc.get("bufferViews");
        size_t bvi = 0;
        if (!as_uint(a->get("bufferView"), (double)arr_len(bvs), bvi)) return fail("accessor without a valid buffer view");
        const JVal* bv = bvs->at(bvi);
        size_t buf = 0, view_off = 0, view_len = 0;
        if (!opt_uint(bv, "buffer", 0, 65536.0, buf) || buf != 0 || !bin) return fail("only the GLB binary chunk is supported as a buffer");
        if (!opt_uint(bv, "byteOffset", 0, lim, view_off) || !opt_uint(bv, "byteLength", 0, lim, view_len))
            return fail("invalid buffer view range");
        if (view_off > bin_len || view_len > bin_len - view_off) return fail("buffer view beyond the BIN chunk");
        const size_t elem = csize * (size_t)ncomp;
        size_t stride = elem;
        if (bv->get("byteStride")) {
            // glTF: a defined byteStride is a multiple of 4 in [4, 252], at least the element size
            if (!opt_uint(bv, "byteStride", 0, 253.0, stride) || stride % 4 || stride < 4 || stride < elem)
                return fail("invalid byteStride");
        }
        if (count) {
            if (acc_off > view_len || elem > view_len - acc_off) return fail("accessor reads beyond its buffer view");
            if (count - 1 > (view_len - acc_off - elem) / stride) return fail("accessor reads beyond its buffer view");
        }
        const size_t off = view_off + acc_off;
        const JVal* nz = a->get("normalized");
        const bool norm = nz && nz->t == JVal::Bool && nz->b;
        if (norm && !allow_norm_int) return fail("normalised accessor where floats are expected");
        if (ct != 5126 && !norm && want_type && std::string(want_type) != "SCALAR") return fail("integer vertex attributes must be normalised");
        out.resize(count * (size_t)ncomp);
        for (size_t j = 0; j < count; ++j)
            for (int c = 0; c < ncomp; ++c) {
                const uint8_t* q = bin + off + j * stride + (size_t)c * csize;
                double v;
                if (ct == 5126) { float f; std::memcpy(&f, q, 4); v = f; }
                else if (ct == 5125) { uint32_t u; std::memcpy(&u, q, 4); v = u; }
                else if (ct == 5123) { uint16_t u; std::memcpy(&u, q, 2); v = u; }
                else { v = q[0]; }
                out[j * ncomp + c] = v;
            }
        return true;
    }

    // first-appearance order of the primitive's base-colour texture, 0 without one
    bool texture_order(const JVal* prim, int& order) {
        order = 0;
        const JVal* mi = prim->get("material");
        if (!mi) return true;
        const JVal* mats = doc.get("materials");
        size_t m_i = 0;
        if (!as_uint(mi, (double)arr_len(mats), m_i)) return fail("material index out of range");
        const JVal* pbr = mats->at(m_i)->get("pbrMetallicRoughness");
        const JVal* bct = pbr ? pbr->get("baseColorTexture") : nullptr;
        if (!bct) return true;
        size_t ti = 0;
        if (!as_uint(bct->get("index"), (double)arr_len(doc.get("textures")), ti)) return fail("texture index out of range");
        if (ti >= tex_of_texture_index.size()) tex_of_texture_index.resize(ti + 1, -1);
        if (tex_of_texture_index[ti] < 0) tex_of_texture_index[ti] = n_textures++;
        order = tex_of_texture_index[ti];
        return true;
    }

    bool primitive(const JVal* prim, const M4& world, Scene& sc) {
        const JVal* attrs = prim->get("attributes");
        const JVal* pa = attrs ? attrs->get("POSITION") : nullptr;
        if (!pa) return fail("primitive without POSITION");
        const JVal* ia = prim->get("indices");
        if (!ia) return fail("non-indexed primitives are not supported");
        size_t mode = 0;
        if (!opt_uint(prim, "mode", 4, 7.0, mode) || mode != 4) return fail("only triangle-list primitives are supported");
        const double n_acc = (double)arr_len(doc.get("accessors"));
        size_t a_pos = 0, a_nor = 0, a_uv = 0, a_idx = 0;
        const JVal* na = attrs->get("NORMAL");
        const JVal* ta = attrs->get("TEXCOORD_0");
        if (!as_uint(pa, n_acc, a_pos) || !as_uint(ia, n_acc, a_idx) || (na && !as_uint(na, n_acc, a_nor)) ||
            (ta && !as_uint(ta, n_acc, a_uv)))
            return fail("accessor index out of range");
        std::vector<double> pos, nor, uv, idx;
        int nc;
        if (!accessor(a_pos, "VEC3", pos, nc, false)) return false;
        const size_t nv = pos.size() / 3;
        if (na && !accessor(a_nor, "VEC3", nor, nc, false)) return false;
        if (ta && !accessor(a_uv, "VEC2", uv, nc, true)) return false;
        if (!nor.empty() && nor.size() != pos.size()) return fail("NORMAL count differs from POSITION count");
        if (!uv.empty() && uv.size() / 2 != nv) return fail("TEXCOORD_0 count differs from POSITION count");
        if (!accessor(a_idx, "SCALAR", idx, nc, false)) return false;
        if (idx.size() % 3) return fail("index count is not a multiple of 3");
        int mat = 0;
        if (!texture_order(prim, mat)) return false;
        const uint32_t voff = (uint32_t)(sc.positions.size() / 3);
        Scene::Object ob;
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        ob.start = (uint32_t)sc.indices.size();
        for (double di : idx) {
            const size_t k = (size_t)di;
            if (k >= nv) return fail("index beyond the vertex count");
            const float p[3] = {(float)pos[3 * k], (float)pos[3 * k + 1], (float)pos[3 * k + 2]};
            double w[3];
            xform_swap(world, p, w);
            for (int a = 0; a < 3; ++a) { mn[a] = std::min(mn[a], w[a]); mx[a] = std::max(mx[a], w[a]); }
            sc.soup.push_back((float)w[0]); sc.soup.push_back((float)w[1]); sc.soup.push_back((float)w[2]); sc.soup.push_back(0.0f);
            sc.indices.push_back((uint32_t)k + voff);
        }
        ob.count = (uint32_t)(sc.indices.size() - ob.start);
        for (int a = 0; a < 3; ++a) { ob.mn[a] = (float)mn[a]; ob.mx[a] = (float)mx[a]; }
        sc.objects.push_back(ob);
        double nm[3][3];
        const bool have_nm = !nor.empty() && normal_matrix(world, nm);
        for (size_t v = 0; v < nv; ++v) {
            const float p[3] = {(float)pos[3 * v], (float)pos[3 * v + 1], (float)pos[3 * v + 2]};
            double w[3];
            xform_swap(world, p, w);
            sc.positions.push_back((float)w[0]); sc.positions.push_back((float)w[1]); sc.positions.push_back((float)w[2]);
            if (!nor.empty()) {
                double n[3] = {0, 0, 0};
                const double x = (float)nor[3 * v], y = (float)nor[3 * v + 1], z = (float)nor[3 * v + 2];
                if (have_nm) for (int r = 0; r < 3; ++r) n[r] = nm[r][0] * x + nm[r][1] * y + nm[r][2] * z;
                double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                if (l == 0.0) l = 1.0;                      // Vector3.normalize divides by (length || 1)
                sc.normals.push_back((float)(n[0] / l)); sc.normals.push_back((float)(n[1] / l)); sc.normals.push_back((float)(n[2] / l));
            }
            sc.uvs.push_back(uv.empty() ? 0.0f : (float)uv[2 * v]);
            sc.uvs.push_back(uv.empty() ? 0.0f : (float)uv[2 * v + 1]);
            sc.material_indices.push_back((uint32_t)mat);
        }
        return true;
    }

    size_t node_visits = 0;

    // glTF nodes form trees: a node is visited at most once, so more visits than nodes (a node
    // listed as a child twice, a cycle) is malformed and stops the walk.
    bool node(const JVal* ref, const M4& parent, Scene& sc, int depth) {
        if (depth > 256) return fail("node hierarchy too deep (cycle?)");
        const JVal* nodes = doc.get("nodes");
        size_t ni = 0;
        if (!as_uint(ref, (double)arr_len(nodes), ni)) return fail("node index out of range");
        if (++node_visits > arr_len(nodes)) return fail("nodes do not form a tree");
        const JVal* n = nodes->at(ni);
        M4 local = ident4();
        if (const JVal* m = n->get("matrix")) {
            if (m->t != JVal::Arr || m->arr.size() != 16) return fail("node matrix must have 16 numbers");
            for (int i = 0; i < 16; ++i) local.e[i] = m->arr[i].num;
        } else {
            double t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
            if (const JVal* v = n->get("translation")) for (int i = 0; i < 3 && v->at(i); ++i) t[i] = v->at(i)->num;
            if (const JVal* v = n->get("rotation")) for (int i = 0; i < 4 && v->at(i); ++i) q[i] = v->at(i)->num;
            if (const JVal* v = n->get("scale")) for (int i = 0; i < 3 && v->at(i); ++i) s[i] = v->at(i)->num;
            local = compose(t, q, s);
        }
        const M4 world = mul4(parent, local);
        if (const JVal* mi = n->get("mesh")) {
            const JVal* meshes = doc.get("meshes");
            size_t m_i = 0;
            if (!as_uint(mi, (double)arr_len(meshes), m_i)) return fail("mesh index out of range");
            const JVal* prims = meshes->at(m_i)->get("primitives");
            if (!prims || prims->t != JVal::Arr) return fail("mesh without primitives");
            for (const JVal& p : prims->arr) if (!primitive(&p, world, sc)) return false;
        }
        if (const JVal* ch = n->get("children"))
            for (const JVal& c : ch->arr) if (!node(&c, world, sc, depth + 1)) return false;
        return true;
    }

    bool load(const uint8_t* glb, size_t len, Scene& sc) {
        auto u32at = [&](size_t o) { uint32_t v; std::memcpy(&v, glb + o, 4); return v; };
        if (len < 20 || u32at(0) != 0x46546C67u) return fail("not a glTF binary (magic)");
        if (u32at(4) != 2u) return fail("unsupported glTF version");
        if (u32at(8) > len) return fail("truncated GLB");
        const size_t total = u32at(8);
        const size_t jlen = u32at(12);
        if (u32at(16) != 0x4E4F534Au || 20 + jlen > total) return fail("first GLB chunk must be JSON");
        size_t o = 20 + jlen;
        if (o + 8 <= total && u32at(o + 4) == 0x004E4942u) {
            const size_t blen = u32at(o);
            if (o + 8 + blen > total) return fail("truncated BIN chunk");
            bin = glb + o + 8;
            bin_len = blen;
        }
        JParser jp{(const char*)glb + 20, (const char*)glb + 20 + jlen, ""};
        if (!jp.value(doc, 0) || doc.t != JVal::Obj) return fail("malformed JSON chunk" + (jp.err.empty() ? std::string() : ": " + jp.err));
        if (const JVal* bufs = doc.get("buffers"))
            for (const JVal& b : bufs->arr) if (b.get("uri")) return fail("external or data-URI buffers are not supported");
        const JVal* scenes = doc.get("scenes");
        size_t si = 0;
        if (!opt_uint(&doc, "scene", 0, 1e9, si) || si >= arr_len(scenes)) return fail("no scene");
        const JVal* s = scenes->at(si);
        if (const JVal* roots = s->get("nodes"))
            for (const JVal& r : roots->arr) if (!node(&r, ident4(), sc, 0)) return false;
        return true;
    }
};

}  // namespace mpm_gltf
