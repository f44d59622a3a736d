// shading.h — device side of the walk's per-vertex work: hit fill and region classification,
// Lambert / smooth dielectric / GGX, quad + point lights with power-CDF selection, the
// procedural sky with its equirect importance table, and the MIS combiner.
// Each function restates the GLSL source cited beside it in the same fp32 operation order.
#pragma once
#include "dev_common.h"
#include "traversal.h"
#include "sky.h"

namespace ptmi {

struct Frame { f3 t, b, n; };
struct Hit {
    float t; f3 p; Frame frame; f3 ng; float eps;
    int owner, region_from, region_to;
};

// fp_uncertainty (glsl/core/math.glsl:52-56)
PT_DEV float fp_uncertainty(f3 p) {
    return gmax(1.52587890625e-5f, 3.0517578125e-5f * gmax(fabsf(p.x), gmax(fabsf(p.y), fabsf(p.z))));
}
// ambient_frame (components/ambient/euclidean/euclidean.glsl:13-19)
PT_DEV Frame ambient_frame(f3 n) {
    Frame f;
    f3 nn = normalize(n);
    f3 t = fabsf(nn.x) < 0.9f ? mk3(1.0f, 0.0f, 0.0f) : mk3(0.0f, 1.0f, 0.0f);
    t = normalize(t - nn * dot(t, nn));
    f.t = t; f.b = cross(nn, t); f.n = nn;
    return f;
}

// ---------------------------------------------------------------- containment
// mesh_inside_bvh / mesh_closest_bvh (intersection/mesh/mesh.glsl:202-296) and
// scene_region_at (features/intersection.ts:1121-1265), innermost wins.
// The inside probe casts a ray along one fixed direction and asks whether the first triangle it
// hits faces away. Every probe ray is parallel, so the triangles it can hit are those whose
// projections onto the plane across the direction cover the ray's point: the probe map's cell
// of that point lists them with their conservatively widened footprints (builder.h); the ones
// whose footprint holds the point are tested, and the closest of their Moller-Trumbore hits under
// the (t, id) rule is the walk's closest hit, with no traversal.
PT_DEV float point_tri_dist2(f3 p, f3 a, f3 b, f3 c) {
    f3 ab = b - a, ac = c - a, ap = p - a;
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) { f3 q = p - a; return dot(q, q); }
    f3 bp = p - b;
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) { f3 q = p - b; return dot(q, q); }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) { f3 q = ap - ab * (d1 / (d1 - d3)); return dot(q, q); }
    f3 cp = p - c;
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) { f3 q = p - c; return dot(q, q); }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) { f3 q = ap - ac * (d2 / (d2 - d6)); return dot(q, q); }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        f3 q = bp - (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
        return dot(q, q);
    }
    float denom = 1.0f / (va + vb + vc);
    f3 closest = (a + ab * (vb * denom)) + ac * (vc * denom);
    f3 q = p - closest;
    return dot(q, q);
}
// The diagonal of an object's box (bounds its points' distances to its triangles).
PT_DEV float box_diagonal(const pt_object& ob) {
    return length(mk3(ob.box_max[0] - ob.box_min[0], ob.box_max[1] - ob.box_min[1], ob.box_max[2] - ob.box_min[2]));
}
// With G > 1 a group of G consecutive lanes answers one probe together (every lane of the group
// calls with the same object and point, lane g of the group testing candidates g, g + G, ...), and
// the group's closest hit is reduced across its lanes: the probe's candidates load in parallel.
template <int G>
PT_DEV void group_closest(float& t, int& id) {
#pragma unroll
    for (int m = 1; m < G; m <<= 1) {
        const float t2 = __shfl_xor(t, m, G);
        const int id2 = __shfl_xor(id, m, G);
        if (t2 < t || (t2 == t && id2 < id)) { t = t2; id = id2; }
    }
}
// on_surface (when given) is set when the probe reports inside and p lies on one of the
// footprint candidates (point_tri_dist2 == 0): a triangle at distance 0 is one of them, since its
// footprint holds p's projection.
template <int G = 1>
PT_DEV bool mesh_inside(const DevScene& S, int obj, f3 p, int g = 0, bool* on_surface = nullptr) {
    const int4 h = S.pm_head[obj];
    if (h.y == 0) return false;
    const float4 gr = S.pm_grid[obj];
    const float u = (S.pm_basis[0] * p.x + S.pm_basis[1] * p.y) + S.pm_basis[2] * p.z;
    const float v = (S.pm_basis[3] * p.x + S.pm_basis[4] * p.y) + S.pm_basis[5] * p.z;
    const float fu = floorf((u - gr.x) * gr.z), fv = floorf((v - gr.y) * gr.w);
    if (!(fu >= 0.0f && fu < (float)h.y && fv >= 0.0f && fv < (float)h.z)) return false;   // no footprint here
    const int c = h.x + (int)fv * h.y + (int)fu;
    const int beg = S.pm_cells[c], end = S.pm_cells[c + 1];
    const f3 dir = mk3(0.5320544f, 0.7396997f, 0.4114743f);
    float tb = 1.0e20f;
    int best = -1;
    for (int k = beg + g; k < end; k += G) {
        const float4 fb = S.pm_boxes[k];                     // contiguous per cell: cheap rejection
        if (u < fb.x || u > fb.y || v < fb.z || v > fb.w) continue;
        const float4* tp = S.ow_tris + 3 * (size_t)S.pm_items[k];
        const float4 ta = tp[0], te1 = tp[1], te2 = tp[2];
        float t;
        if (tri_test(p, dir, xyz(ta), xyz(te1), xyz(te2), t) && t > 0.0f) {
            const int id = (int)__float_as_uint(ta.w);
            if (t < tb || (t == tb && id < best)) { tb = t; best = id; }
        }
    }
    if (G > 1) group_closest<G>(tb, best);
    if (best < 0) return false;
    const float4* tp = S.tri_shade + 3 * (size_t)best;
    f3 a = xyz(tp[0]);
    const bool inside = dot(dir, cross(xyz(tp[1]) - a, xyz(tp[2]) - a)) > 0.0f;
    if (inside && on_surface) {
        bool on = false;
        for (int k = beg + g; k < end; k += G) {
            const float4 fb = S.pm_boxes[k];
            if (u < fb.x || u > fb.y || v < fb.z || v > fb.w) continue;
            const int id = (int)__float_as_uint(S.ow_tris[3 * (size_t)S.pm_items[k]].w);
            const float4* sp = S.tri_shade + 3 * (size_t)id;
            on = on || point_tri_dist2(p, xyz(sp[0]), xyz(sp[1]), xyz(sp[2])) == 0.0f;
        }
        if (G > 1) {
#pragma unroll
            for (int m = 1; m < G; m <<= 1) on = on || __shfl_xor((int)on, m, G) != 0;
        }
        *on_surface = on;
    }
    return inside;
}
PT_DEV float box_dist2(float4 n0, float4 n1, f3 p) {
    f3 d = mk3(gmax(gmax(n0.x - p.x, 0.0f), p.x - n1.x), gmax(gmax(n0.y - p.y, 0.0f), p.y - n1.y),
               gmax(gmax(n0.z - p.z, 0.0f), p.z - n1.z));
    return dot(d, d);
}
PT_DEV float mesh_closest(const DevScene& S, int obj, f3 p) {
    const int base = S.obj_node_base[obj];
    if (base < 0) return 1.0e20f;
    float best2 = 1.0e30f;
    int stack[64]; int sp = 0; stack[0] = 0;
    while (sp >= 0) {
        int ni = stack[sp--];
        float4 n0 = S.obj_nodes[2 * (base + ni)], n1 = S.obj_nodes[2 * (base + ni) + 1];
        if (box_dist2(n0, n1, p) >= best2) continue;
        if (n0.w >= 0.0f) {
            int off = (int)n1.w, cnt = (int)n0.w;
            const pt_object& ob = S.objects[obj];
            int first = ob.kind == PT_OBJ_INSTANCES ? S.n_tris + ob.first_tri : ob.first_tri;
            for (int i = off; i < off + cnt; i++) {
                int tri = S.obj_order[first + i];
                const float4* tp = S.tri_shade + 3 * (size_t)tri;
                best2 = gmin(best2, point_tri_dist2(p, xyz(tp[0]), xyz(tp[1]), xyz(tp[2])));
            }
        } else {
            int L = ni + 1, R = (int)n1.w;
            bool left_near = box_dist2(S.obj_nodes[2 * (base + L)], S.obj_nodes[2 * (base + L) + 1], p) <=
                             box_dist2(S.obj_nodes[2 * (base + R)], S.obj_nodes[2 * (base + R) + 1], p);
            if (sp + 2 < 64) { stack[++sp] = left_near ? R : L; stack[++sp] = left_near ? L : R; }
        }
    }
    return sqrtf(best2);
}
// scene_region_at: every closed object holding p has depth d = -(distance to its nearest
// triangle) (sphere: |p - c| - r), and the region is the deepest one, the first in the
// reference's order on ties (spheres, then the grid's meshes, then the instance batches). It runs
// in two passes. The first only asks which objects hold p (sphere test, box test and inside
// probe) and, when exactly one does, returns it: the depths only order several holders. That
// holder's depth is strictly negative and above the -1e20 floor the comparison starts from, which
// the first pass makes sure of: a mesh's distance is at most its box's diagonal times the
// placement scale (checked against 1e19), and it is 0 only if p lies on one of the triangles,
// which then is one of the probe's footprint candidates (mesh_inside reports it). Otherwise, and
// with two or more holders, the second pass computes the depths (the closest-triangle search).
template <int G = 1, bool DEPTHS>
PT_DEV int region_pass(const DevScene& S, f3 p, int g, int& holders) {
    int region = -1; float best = -1.0e20f;
    holders = 0;
    for (int k = 0; k < S.n_analytic; k++) {
        int o = S.analytic[k];
        const pt_object& ob = S.objects[o];
        if (ob.kind != PT_OBJ_SPHERE) continue;
        float d = length(p - ld3(ob.p)) - ob.p[3];
        if (d < 0.0f && d > best) {
            if (DEPTHS) best = d;
            region = o;
            holders++;
        }
    }
    if (S.has_grid) {
        int c[3];
        const float pc[3] = {p.x, p.y, p.z};
        bool inside_grid = true;
        for (int a = 0; a < 3; a++) {
            float f = floorf((pc[a] - S.grid_lo[a]) * S.grid_inv[a]);
            if (!(f >= 0.0f && f < (float)S.grid_n[a])) inside_grid = false;   // outside every closed-mesh box
            c[a] = inside_grid ? (int)f : 0;
        }
        if (inside_grid) {
            int2 cell = S.grid_cells[(c[2] * S.grid_n[1] + c[1]) * S.grid_n[0] + c[0]];
            for (int j = 0; j < cell.y; j++) {
                int o = S.grid_items[cell.x + j];
                const pt_object& ob = S.objects[o];
                bool on_surface = false;
                if (p.x >= ob.box_min[0] && p.y >= ob.box_min[1] && p.z >= ob.box_min[2] &&
                    p.x <= ob.box_max[0] && p.y <= ob.box_max[1] && p.z <= ob.box_max[2] &&
                    mesh_inside<G>(S, o, p, g, &on_surface)) {
                    if (DEPTHS) {
                        const float d = -1.0f * mesh_closest(S, o, p);
                        if (d < 0.0f && d > best) { best = d; region = o; }
                    } else {
                        region = o;
                        holders += on_surface || !(box_diagonal(ob) < 1e19f) ? 2 : 1;
                    }
                }
            }
        }
    }
    // instanced batches claiming an interior (features/intersection.ts:1196-1215): point descent
    // of the instance TLAS, the prototype's frame (/ s), world distance s x closest distance
    if (S.n_inst_nodes > 0) {
        int stack[64]; int sp = 0; stack[0] = 0;
        while (sp >= 0) {
            const int ni = stack[sp--];
            const float4 n0 = S.inst_nodes[2 * ni], n1 = S.inst_nodes[2 * ni + 1];
            if (!(p.x >= n0.x && p.y >= n0.y && p.z >= n0.z && p.x <= n1.x && p.y <= n1.y && p.z <= n1.z)) continue;
            if (n0.w >= 0.0f) {
                const int off = (int)n1.w, count = (int)n0.w;
                for (int j = off; j < off + count; j++) {
                    const int i = S.inst_order[j];
                    const int o = S.inst_batch[i];
                    const pt_object& ob = S.objects[o];
                    if (!ob.closed) continue;
                    const float4 q = S.inst_q[i], ts = S.inst_ts[i];
                    const f3 r = qrot(q, p) + xyz(ts);
                    const f3 lp = mk3(r.x / ts.w, r.y / ts.w, r.z / ts.w);
                    bool on_surface = false;
                    if (lp.x >= ob.box_min[0] && lp.y >= ob.box_min[1] && lp.z >= ob.box_min[2] &&
                        lp.x <= ob.box_max[0] && lp.y <= ob.box_max[1] && lp.z <= ob.box_max[2] &&
                        mesh_inside<G>(S, o, lp, g, &on_surface)) {
                        if (DEPTHS) {
                            const float d = -ts.w * mesh_closest(S, o, lp);
                            if (d < 0.0f && d > best) { best = d; region = o; }
                        } else {
                            region = o;
                            holders += on_surface || !(ts.w * box_diagonal(ob) < 1e19f) ? 2 : 1;
                        }
                    }
                }
            } else if (sp + 2 < 64) {
                stack[++sp] = ni + 1; stack[++sp] = (int)n1.w;
            }
        }
    }
    return region;
}
template <int G = 1>
PT_DEV int region_at(const DevScene& S, f3 p, int g = 0) {
    int holders;
    const int region = region_pass<G, false>(S, p, g, holders);
    if (holders <= 1) return region;
    return region_pass<G, true>(S, p, g, holders);
}
PT_DEV bool object_thin(const DevScene& S, int o) {
    const pt_object& ob = S.objects[o];
    return ob.kind == PT_OBJ_QUAD || ((ob.kind == PT_OBJ_MESH || ob.kind == PT_OBJ_INSTANCES) && !ob.closed);
}
// The path's four wavelengths (nm); lane 0 is the hero (spectral programs only).
struct Wl { float4 lam; };
// Regions with a row in the generated ior_of table (features/intersection.ts:1609-1649): every
// primitive object, closed meshes and closed instance batches; open meshes are thin (1.0).
// The models that cross their interface and read the region's index (the descriptors'
// 'transmission' capability: dielectric and rough_dielectric).
__host__ __device__ inline bool transmissive_model(int model) { return model == PT_MAT_DIELECTRIC || model == PT_MAT_ROUGH_DIELECTRIC; }
PT_DEV bool region_has_index_row(const pt_object& ob) {
    return !((ob.kind == PT_OBJ_MESH || ob.kind == PT_OBJ_INSTANCES) && !ob.closed);
}
// ior_of: a dielectric region's index; a dispersive formula under the spectral colour model is
// read at the hero wavelength (spectral.glsl:12-15; cauchy.glsl, sellmeier.glsl, micrometres).
template <bool SPEC>
PT_DEV float ior_of(const DevScene& S, int region, const Wl& wl) {
    if (region < 0) return 1.0f;
    const pt_object& ob = S.objects[region];
    if (!region_has_index_row(ob)) return 1.0f;
    const pt_material& m = S.materials[ob.material];
    if (!transmissive_model(m.model)) return 1.0f;
    if (SPEC && m.ior_kind == PT_IOR_CAUCHY) {
        const float lam = wl.lam.x;
        const float inv2 = 1e6f / (lam * lam);
        return m.ior_coef[0] + inv2 * (m.ior_coef[1] + inv2 * m.ior_coef[2]);
    }
    if (SPEC && m.ior_kind == PT_IOR_SELLMEIER) {
        const float lam = wl.lam.x;
        const float l2 = (lam * 0.001f) * (lam * 0.001f);
        const float t0 = m.ior_coef[0] * l2 / (l2 - m.ior_coef[3]), t1 = m.ior_coef[1] * l2 / (l2 - m.ior_coef[4]),
                    t2 = m.ior_coef[2] * l2 / (l2 - m.ior_coef[5]);
        return sqrtf(1.0f + t0 + t1 + t2);
    }
    return m.ior;
}
// A region of the program's dispersion record (Planner.ts:1038-1053 planDispersion).
PT_DEV bool dispersive_region(const DevScene& S, int region) {
    if (region < 0) return false;
    const pt_object& ob = S.objects[region];
    const pt_material& m = S.materials[ob.material];
    return region_has_index_row(ob) && transmissive_model(m.model) && m.ior_kind != PT_IOR_CONSTANT;
}

// The placement owning an instance primitive id: the last instance whose first id <= prim.
PT_DEV int instance_of(const DevScene& S, int prim) {
    const int rel = prim - S.inst_prim0;
    int lo = 0, hi = S.n_instances - 1;
    while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (S.inst_base[mid] <= rel) lo = mid; else hi = mid - 1; }
    return lo;
}
// The region object of a primitive id: triangle owner, instance batch or analytic object.
PT_DEV int prim_owner(const DevScene& S, int prim) {
    if (prim < S.n_tris) return S.tri_object[prim];
    if (prim >= S.inst_prim0 && S.n_instances > 0) return S.inst_batch[instance_of(S, prim)];
    return prim - S.n_tris;
}

// Hit fill (features/intersection.ts:889-948 mesh wrapper, primitive arms) and the
// dispatcher's orientation step (features/intersection.ts:1690-1706). The classification
// probes run only when something reads region_from / region_to: a transmissive material
// (ior_of) or an emissive solid region (emission keys on region_to).
// mesh_test_range's smooth shading normal (mesh.glsl:84-109): the hit's barycentrics from
// mesh_tri_test, bary.x n_a + bary.y n_b + bary.z n_c, turned to the geometric normal's ray
// side; (o, d) is the ray in the mesh's frame, tri the shade-triangle index.
PT_DEV f3 smooth_normal(const DevScene& S, int tri, f3 o, f3 d, f3 a, f3 b, f3 c) {
    const f3 e1 = b - a, e2 = c - a;
    const f3 g = cross(e1, e2);
    const float invdet = 1.0f / -dot(d, g);
    const f3 ao = o - a;
    const f3 dao = cross(ao, d);
    const float u = dot(e2, dao) * invdet;
    const float v = -dot(e1, dao) * invdet;
    const float w = 1.0f - u - v;
    const float* n = S.tri_nrm + 9 * (size_t)tri;
    f3 ns = (mk3(n[0], n[1], n[2]) * w + mk3(n[3], n[4], n[5]) * u) + mk3(n[6], n[7], n[8]) * v;
    if (dot(d, ns) * dot(d, g) < 0.0f) ns = -ns;
    return ns;
}
// The hit point, the unflipped geometric normal and the owning region of a closest hit: the part
// of the hit fill the region probes read.
PT_DEV void hit_geometry(const DevScene& S, f3 o, f3 d, float t, int prim, f3& p, f3& ng, int& owner) {
    p = o + d * t;
    if (prim < S.n_tris) {
        const float4* tp = S.tri_shade + 3 * (size_t)prim;
        const f3 a = xyz(tp[0]);
        ng = normalize(cross(xyz(tp[1]) - a, xyz(tp[2]) - a));
        owner = S.tri_object[prim];
    } else if (prim >= S.inst_prim0 && S.n_instances > 0) {
        // instanceLeafItem mesh arm: the prototype triangle's normal by the forward rotation
        const int rel = prim - S.inst_prim0;
        const int lo = instance_of(S, prim);
        const int batch = S.inst_batch[lo];
        const int tri = S.n_tris + S.objects[batch].first_tri + (rel - S.inst_base[lo]);
        const float4* tp = S.tri_shade + 3 * (size_t)tri;
        const f3 a = xyz(tp[0]);
        const float4 q = S.inst_q[lo];
        ng = normalize(qrot(make_float4(-q.x, -q.y, -q.z, q.w), cross(xyz(tp[1]) - a, xyz(tp[2]) - a)));
        owner = batch;
    } else {
        const int o2 = prim - S.n_tris;
        const pt_object& ob = S.objects[o2];
        ng = normalize(ob.kind == PT_OBJ_QUAD ? ld3(ob.p + 9) : normalize(p - ld3(ob.p)));
        owner = o2;
    }
}
PT_DEV void fill_hit(const DevScene& S, f3 o, f3 d, float t, int prim, Hit& h, bool& back) {
    h.t = t;
    hit_geometry(S, o, d, t, prim, h.p, h.ng, h.owner);
    h.eps = fp_uncertainty(h.p);
    if (prim < S.n_tris) {
        const float4* tp = S.tri_shade + 3 * (size_t)prim;
        f3 a = xyz(tp[0]);
        f3 g = cross(xyz(tp[1]) - a, xyz(tp[2]) - a);
        h.frame = ambient_frame(normalize(S.objects[h.owner].smooth ? smooth_normal(S, prim, o, d, a, xyz(tp[1]), xyz(tp[2])) : g));
    } else if (prim >= S.inst_prim0 && S.n_instances > 0) {
        f3 ns = h.ng;
        if (S.objects[h.owner].smooth) {   // the placement's local ray, as the prototype walk saw it
            const int rel = prim - S.inst_prim0;
            const int lo = instance_of(S, prim);
            const int tri = S.n_tris + S.objects[h.owner].first_tri + (rel - S.inst_base[lo]);
            const float4* tp = S.tri_shade + 3 * (size_t)tri;
            const f3 a = xyz(tp[0]);
            const float4 q = S.inst_q[lo];
            const float4 qc = make_float4(-q.x, -q.y, -q.z, q.w);
            const float4 ts = S.inst_ts[lo];
            const f3 r = qrot(q, o) + xyz(ts), rd = qrot(q, d);
            const f3 lo3 = mk3(r.x / ts.w, r.y / ts.w, r.z / ts.w), ld = mk3(rd.x / ts.w, rd.y / ts.w, rd.z / ts.w);
            ns = normalize(qrot(qc, smooth_normal(S, tri, lo3, ld, a, xyz(tp[1]), xyz(tp[2]))));
        }
        h.frame = ambient_frame(ns);
    } else {
        const pt_object& ob = S.objects[h.owner];
        // ambient_frame normalises its argument, so frame.n is the normal hit_geometry returned
        h.frame = ambient_frame(ob.kind == PT_OBJ_QUAD ? ld3(ob.p + 9) : normalize(h.p - ld3(ob.p)));
    }
    back = !(dot(d, h.ng) < 0.0f);
}
template <int G = 1>
PT_DEV void classify_regions(const DevScene& S, Hit& h, bool back, bool need, int g = 0) {
    // h.ng is still the unflipped geometric normal here
    if (need) {
        int outside = region_at<G>(S, h.p + h.ng * EPS_INTERFACE, g);
        if (!back) { h.region_from = outside; h.region_to = h.owner; }
        else {
            h.region_from = object_thin(S, h.owner) ? region_at<G>(S, h.p + h.ng * (-EPS_INTERFACE), g) : h.owner;
            h.region_to = outside;
        }
    } else {
        h.region_from = -2;
        h.region_to = back ? -2 : h.owner;   // -2: not read (no emissive solid can be behind)
    }
}
// The dispatcher's orientation step: a back-face hit turns the frame and normal to the ray's side.
PT_DEV void orient_hit(Hit& h, bool back) {
    if (back) {
        h.frame = ambient_frame(-h.frame.n);
        h.ng = -h.ng;
    }
}
PT_DEV void classify_hit(const DevScene& S, Hit& h, bool back, bool need) {
    classify_regions(S, h, back, need);
    orient_hit(h, back);
}

// ---------------------------------------------------------------- spectra
// Wavelengths of a path (spectral programs): four hero lanes.
PT_DEV float glsl_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
// spectral_curves (components/color/spectral/lookup.glsl:2-7): linear interpolation of the
// 5 nm rows, int() truncation of the row coordinate.
PT_DEV float4 spectral_curves(const DevScene& S, float lam) {
    float x = (lam - 360.0f) / 5.0f;
    int i = (int)x; i = i < 93 ? i : 93;
    float f = x - (float)i;
    float4 a = S.spec_curves[i], b = S.spec_curves[i + 1];
    return make_float4(glsl_mix(a.x, b.x, f), glsl_mix(a.y, b.y, f), glsl_mix(a.z, b.z, f), glsl_mix(a.w, b.w, f));
}
PT_DEV float lane(const float4& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }
PT_DEV float spectral_bounded_lane(const pt_spectrum& r, float lam) {   // lookup.glsl:40-48
    float t = (lam - 360.0f) / 470.0f;
    float p = (r.c[0] * t + r.c[1]) * t + r.c[2];
    return gclamp(0.5f + 0.5f * p / sqrtf(1.0f + p * p), 0.0f, 1.0f);
}
// blackbody_emission (color/spectra/blackbody/blackbody.glsl:4-10); 14387768.775039338 is
// the GLSL literal, rounded to fp32.
PT_DEV float blackbody_lane(float kelvin, float normalization, float lam) {
    float q = 14387768.775039338f / kelvin;
    float ratio = 550.0f / lam;
    float ratio2 = ratio * ratio;
    return normalization * (ptm_exp(q / 550.0f) - 1.0f) * ratio2 * ratio2 * ratio / (ptm_exp(q / lam) - 1.0f);
}
// A pt_spectrum at the path's four wavelengths: color_lift_{reflectance,coefficient,emission}
// (lookup.glsl:50-60) for an authored RGB, scale * blackbody_emission, or a Gaussian band.
PT_DEV sp4 spectrum_eval(const DevScene& S, const pt_spectrum& r, const Wl& wl) {
    float v[4];
    for (int i = 0; i < 4; i++) {
        const float lam = lane(wl.lam, i);
        float x;
        if (r.kind == PT_SPEC_BLACKBODY) {
            x = r.scale * blackbody_lane(r.c[0], r.c[1], lam);
        } else if (r.kind == PT_SPEC_GAUSSIAN) {          // gaussian.ts emit: floor + A exp(B s s)
            const float sh = lam - r.c[0];
            x = r.value + r.c[2] * ptm_exp((r.c[1] * sh) * sh);
        } else {
            x = r.kind == PT_SPEC_ZERO ? 0.0f : (r.kind == PT_SPEC_GREY ? r.value : spectral_bounded_lane(r, lam));
            if (r.scaled) x = r.scale * x;
            if (r.d65) x = x * spectral_curves(S, lam).w;
        }
        v[i] = x;
    }
    return mk4(v[0], v[1], v[2], v[3]);
}
// The color model's lift of a material / light property: the authored RGB in an RGB program,
// the lifted spectrum in a spectral one.
template <bool SPEC>
PT_DEV sp4 lift(const DevScene& S, const float* rgb, const pt_spectrum& r, const Wl& wl) {
    if (!SPEC) return rgb4(rgb);
    return spectrum_eval(S, r, wl);
}
// wavelength_from_cdf / wavelength_pdf (color/spectral/visible.glsl:3-13, uniform.glsl:1-2)
PT_DEV float wl_visible_a() { return ptm_tanh(0.0072f * (360.0f - 538.0f)); }
PT_DEV float wl_visible_b() { return ptm_tanh(0.0072f * (830.0f - 538.0f)); }
PT_DEV float wavelength_from_cdf(bool uniform, float u) {
    if (uniform) return 360.0f + 470.0f * u;
    return 538.0f + ptm_atanh(glsl_mix(wl_visible_a(), wl_visible_b(), u)) / 0.0072f;
}
PT_DEV float wavelength_pdf(bool uniform, float lam) {
    if (uniform) return 1.0f / 470.0f;
    float c = ptm_cosh(0.0072f * (lam - 538.0f));
    return 0.0072f / ((wl_visible_b() - wl_visible_a()) * c * c);
}
// sensor_response (sensor/cie1931/cie1931.glsl:3-5): CIE_XYZ_TO_RGB * xyz / CIE_Y_INTEGRAL
PT_DEV f3 sensor_response(const DevScene& S, float lam) {
    float4 c = spectral_curves(S, lam);
    const float* m = S.xyz_rgb;
    f3 r = mk3((m[0] * c.x + m[1] * c.y) + m[2] * c.z, (m[3] * c.x + m[4] * c.y) + m[5] * c.z,
               (m[6] * c.x + m[7] * c.y) + m[8] * c.z);
    return r / S.cie_y;
}

// ---------------------------------------------------------------- materials
enum { LOBE_REFLECTION = 1, LOBE_TRANSMISSION = 2, LOBE_DELTA = 4 };
struct ISample { f3 wi; sp4 weight; float pdf; int flags; };

PT_DEV f3 to_local(const Frame& f, f3 v) { return mk3(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
PT_DEV f3 from_local(const Frame& f, f3 v) { return normalize((f.t * v.x + f.b * v.y) + f.n * v.z); }
// microfacet.glsl:29-54
PT_DEV float mf_D(f3 h, float a) { float t = h.z * h.z * (a * a - 1.0f) + 1.0f; return a * a / (PTM_PI * t * t); }
PT_DEV float mf_G1(f3 v, float a) { float c = fabsf(v.z); return 2.0f * c / (c + sqrtf(a * a + (1.0f - a * a) * c * c)); }
PT_DEV f3 mf_vndf(f3 wol, float a, float ux, float uy) {
    f3 vh = normalize(mk3(a * wol.x, a * wol.y, wol.z));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    f3 T1 = lensq > 0.0f ? mk3(-vh.y, vh.x, 0.0f) * (1.0f / sqrtf(lensq)) : mk3(1.0f, 0.0f, 0.0f);
    f3 T2 = cross(vh, T1);
    float rr = sqrtf(ux);
    float phi = PTM_TWO_PI * uy;
    float sn, cs; ptm_sincos(phi, &sn, &cs);
    float t1 = rr * cs;
    float t2 = rr * sn;
    float s_ = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - s_) * sqrtf(gmax(0.0f, 1.0f - t1 * t1)) + s_ * t2;
    f3 nh = (T1 * t1 + T2 * t2) + vh * sqrtf(gmax(0.0f, 1.0f - t1 * t1 - t2 * t2));
    return normalize(mk3(a * nh.x, a * nh.y, gmax(1e-6f, nh.z)));
}
PT_DEV sp4 schlick(sp4 f0, float c) {
    float p = ptm_pow(1.0f - c, 5.0f);
    return mk4(f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p, f0.w + (1.0f - f0.w) * p);
}
PT_DEV float dielectric_fresnel(float cos_i, float eta) {   // math.glsl:95-102
    float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = sqrtf(1.0f - sin2_t);
    float r_par = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    float r_perp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    return 0.5f * (r_par * r_par + r_perp * r_perp);
}
// MODEL < 0: dispatch on m.model at run time; otherwise the material is known to be MODEL
// (category kernels shade one model per launch).
// (the checker is the Lambert lobe with a per-hit albedo)
template <int MODEL> PT_DEV bool is_model(const pt_material& m, int x) {
    return MODEL < 0 ? (m.model == x || (x == PT_MAT_LAMBERT && m.model == PT_MAT_CHECKER)) : MODEL == x;
}

// materials/rough_dielectric/rough_dielectric.glsl:37-46: the generalized half vector of a
// (wi, wo) pair, faced up; false for a degenerate pair or a back-facing microfacet.
PT_DEV bool rough_half(f3 wil, f3 wol, float eta, f3& m) {
    float etap = wil.z > 0.0f ? 1.0f : 1.0f / eta;
    m = wil * etap + wol;
    if (dot(m, m) <= 0.0f) return false;
    m = normalize(m);
    if (m.z < 0.0f) m = -m;
    return dot(m, wil) * wil.z >= 0.0f && dot(m, wol) * wol.z >= 0.0f;
}

// Hit.uv (glsl/core/structs.glsl:40-43): the owning shape's chart, quad_uv (quad.glsl:33-41) or
// sphere_uv (sphere.glsl:84-89), else the planar placeholder (x, z) * UV_PLANAR_SCALE (0.1).
PT_DEV float2 hit_uv(const DevScene& S, const Hit& h) {
    const pt_object& o = S.objects[h.owner];
    if (o.kind == PT_OBJ_QUAD) {
        const f3 local = h.p - mk3(o.p[0], o.p[1], o.p[2]);
        const f3 e1 = mk3(o.p[3], o.p[4], o.p[5]), e2 = mk3(o.p[6], o.p[7], o.p[8]);
        const float e11 = dot(e1, e1), e22 = dot(e2, e2), e12 = dot(e1, e2);
        const float d1 = dot(local, e1), d2 = dot(local, e2);
        const float det = e11 * e22 - e12 * e12;
        return make_float2((d1 * e22 - d2 * e12) / det, (d2 * e11 - d1 * e12) / det);
    }
    if (o.kind == PT_OBJ_SPHERE) {
        const f3 d = normalize(h.p - mk3(o.p[0], o.p[1], o.p[2]));
        return make_float2(ptm_atan2(d.z, d.x) / PTM_TWO_PI + 0.5f, ptm_acos(gclamp(d.y, -1.0f, 1.0f)) / PTM_PI);
    }
    return make_float2(h.p.x * 0.1f, h.p.z * 0.1f);
}
// The Lambert lobe's albedo: the material's, or the checker's cell (checker.glsl:13-17:
// mix(albedo_a, albedo_b, mod(floor(uv.x s) + floor(uv.y s), 2))).
template <bool SPEC>
PT_DEV sp4 lambert_albedo(const DevScene& S, const pt_material& m, const Hit& h, const Wl& wl) {
    const sp4 a = lift<SPEC>(S, m.albedo, m.s_albedo, wl);
    if (m.model != PT_MAT_CHECKER) return a;
    const float2 uv = hit_uv(S, h);
    const float x = floorf(uv.x * m.pad1) + floorf(uv.y * m.pad1);
    const float k = x - 2.0f * floorf(x / 2.0f);
    return a * (1.0f - k) + lift<SPEC>(S, m.f0, m.s_f0, wl) * k;
}

template <int MODEL, bool SPEC>
PT_DEV ISample surface_sample(const DevScene& S, const pt_material& m, f3 wo, const Hit& h, float uc, float ux, float uy, const Wl& wl) {
    ISample s;
    const Frame& f = h.frame;
    if (is_model<MODEL>(m, PT_MAT_LAMBERT)) {            // materials/lambert/lambert.glsl:17-28
        f3 n = dot(wo, f.n) < 0.0f ? -f.n : f.n;
        float cos_theta = sqrtf(uy);
        float sin_theta = sqrtf(gmax(0.0f, 1.0f - uy));
        float phi = PTM_TWO_PI * ux;
        float sn, cs; ptm_sincos(phi, &sn, &cs);
        s.wi = normalize((f.t * (sin_theta * cs) + f.b * (sin_theta * sn)) + n * cos_theta);
        s.weight = lambert_albedo<SPEC>(S, m, h, wl);
        s.pdf = cos_theta * INV_PI;
        s.flags = LOBE_REFLECTION;
    } else if (is_model<MODEL>(m, PT_MAT_DIELECTRIC)) {  // materials/dielectric/dielectric.glsl:14-52
        float n_i = ior_of<SPEC>(S, h.region_from, wl);
        float n_t = ior_of<SPEC>(S, h.region_to, wl);
        float eta = n_i / n_t;
        f3 n = f.n;
        float cos_i = gclamp(dot(wo, n), 1e-6f, 1.0f);
        float F = dielectric_fresnel(cos_i, eta);
        if (uc < F) {
            s.wi = normalize(n * (2.0f * cos_i) - wo);
            s.weight = sgrey(1.0f);
            s.pdf = 0.0f;
            s.flags = LOBE_REFLECTION | LOBE_DELTA;
        } else {
            float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
            float cos_t = sqrtf(gmax(0.0f, 1.0f - sin2_t));
            s.wi = normalize(wo * (-eta) + n * (eta * cos_i - cos_t));
            float radiance_scale = (n_i * n_i) / (n_t * n_t);
            s.weight = lift<SPEC>(S, m.transmittance, m.s_transmittance, wl) * radiance_scale;
            s.pdf = 0.0f;
            s.flags = LOBE_TRANSMISSION | LOBE_DELTA;
        }
    } else if (is_model<MODEL>(m, PT_MAT_MIRROR)) {      // materials/mirror/mirror.glsl:21-33
        f3 n = f.n;
        float cos_i = gclamp(dot(wo, n), 1e-6f, 1.0f);
        s.wi = normalize(n * (2.0f * cos_i) - wo);
        s.weight = schlick(lift<SPEC>(S, m.f0, m.s_f0, wl), cos_i);
        s.pdf = 0.0f;
        s.flags = LOBE_REFLECTION | LOBE_DELTA;
    } else if (is_model<MODEL>(m, PT_MAT_ROUGH_DIELECTRIC)) {   // rough_dielectric.glsl:74-133
        f3 wol = to_local(f, wo);
        s.wi = wo; s.weight = sgrey(0.0f); s.pdf = 0.0f; s.flags = LOBE_REFLECTION;
        if (wol.z <= 0.0f) return s;
        float a = m.alpha;
        float n_i = ior_of<SPEC>(S, h.region_from, wl);
        float n_t = ior_of<SPEC>(S, h.region_to, wl);
        float eta = n_i / n_t;
        if (n_i == n_t) {                           // index-matched: the delta pass-through
            s.wi = -wo;
            s.weight = lift<SPEC>(S, m.transmittance, m.s_transmittance, wl);
            s.pdf = 0.0f;
            s.flags = LOBE_TRANSMISSION | LOBE_DELTA;
            return s;
        }
        f3 mh = mf_vndf(wol, a, ux, uy);
        float cos_om = gclamp(dot(wol, mh), 1e-6f, 1.0f);
        float F = dielectric_fresnel(cos_om, eta);
        f3 wil;
        if (uc < F) {
            f3 I = -wol;
            wil = I - mh * (2.0f * dot(mh, I));
            if (wil.z <= 0.0f) return s;
            s.weight = sgrey(mf_G1(wil, a));
            s.pdf = F * mf_G1(wol, a) * mf_D(mh, a) / (4.0f * wol.z);
            s.flags = LOBE_REFLECTION;
        } else {
            float sin2_t = eta * eta * (1.0f - cos_om * cos_om);
            float cos_t = sqrtf(gmax(0.0f, 1.0f - sin2_t));
            wil = normalize(wol * (-eta) + mh * (eta * cos_om - cos_t));
            if (wil.z >= 0.0f) return s;
            float denom = dot(wil, mh) + dot(wol, mh) * eta;
            denom *= denom;
            s.weight = lift<SPEC>(S, m.transmittance, m.s_transmittance, wl) * (mf_G1(wil, a) * eta * eta);
            s.pdf = (1.0f - F) * mf_G1(wol, a) * mf_D(mh, a) * cos_om / wol.z * fabsf(dot(wil, mh)) / denom;
            s.flags = LOBE_TRANSMISSION;
        }
        s.wi = from_local(f, wil);
    } else {                                    // materials/ggx/ggx.glsl:27-46
        f3 wol = to_local(f, wo);
        float side = wol.z < 0.0f ? -1.0f : 1.0f;
        wol = wol * side;
        float a = m.alpha;
        f3 hh = mf_vndf(wol, a, ux, uy);
        f3 I = -wol;
        f3 wil = I - hh * (2.0f * dot(hh, I));
        if (wil.z <= 0.0f) {
            s.wi = wo; s.weight = sgrey(0.0f); s.pdf = 0.0f; s.flags = LOBE_REFLECTION;
            return s;
        }
        sp4 F = schlick(lift<SPEC>(S, m.f0, m.s_f0, wl), fabsf(dot(wol, hh)));
        s.weight = F * mf_G1(wil, a);
        s.pdf = mf_G1(wol, a) * mf_D(hh, a) / (4.0f * fabsf(wol.z));
        s.flags = LOBE_REFLECTION;
        s.wi = from_local(f, wil * side);
    }
    return s;
}
template <int MODEL, bool SPEC>
PT_DEV sp4 surface_eval(const DevScene& S, const pt_material& m, f3 wi, f3 wo, const Hit& h, const Wl& wl) {
    const Frame& f = h.frame;
    if (is_model<MODEL>(m, PT_MAT_LAMBERT)) {
        if (dot(wi, f.n) * dot(wo, f.n) <= 0.0f) return sgrey(0.0f);
        return lambert_albedo<SPEC>(S, m, h, wl) * INV_PI;
    } else if (is_model<MODEL>(m, PT_MAT_GGX)) {
        f3 wil = to_local(f, wi), wol = to_local(f, wo);
        if (wil.z * wol.z <= 0.0f) return sgrey(0.0f);
        float a = m.alpha;
        f3 hh = normalize(wil + wol);
        sp4 F = schlick(lift<SPEC>(S, m.f0, m.s_f0, wl), fabsf(dot(wol, hh)));
        return F * (mf_D(hh, a) * mf_G1(wil, a) * mf_G1(wol, a) / (4.0f * fabsf(wil.z) * fabsf(wol.z)));
    } else if (is_model<MODEL>(m, PT_MAT_ROUGH_DIELECTRIC)) {   // rough_dielectric.glsl:48-72
        f3 wil = to_local(f, wi), wol = to_local(f, wo);
        if (wol.z <= 0.0f || wil.z == 0.0f) return sgrey(0.0f);
        float n_i = ior_of<SPEC>(S, h.region_from, wl);
        float n_t = ior_of<SPEC>(S, h.region_to, wl);
        if (n_i == n_t) return sgrey(0.0f);
        float eta = n_i / n_t;
        f3 mh;
        if (!rough_half(wil, wol, eta, mh)) return sgrey(0.0f);
        float a = m.alpha;
        float cos_om = gclamp(fabsf(dot(wol, mh)), 1e-6f, 1.0f);
        float F = dielectric_fresnel(cos_om, eta);
        float DG = mf_D(mh, a) * mf_G1(wol, a) * mf_G1(wil, a);
        if (wil.z > 0.0f) return sgrey(DG * F / (4.0f * wil.z * wol.z));
        float denom = dot(wil, mh) + dot(wol, mh) * eta;
        denom *= denom;
        float ft = DG * (1.0f - F) * fabsf(dot(wil, mh) * dot(wol, mh) / (wil.z * wol.z * denom));
        return lift<SPEC>(S, m.transmittance, m.s_transmittance, wl) * (ft * eta * eta);
    }
    return sgrey(0.0f);
}
template <int MODEL, bool SPEC>
PT_DEV float surface_pdf(const DevScene& S, const pt_material& m, f3 wi, f3 wo, const Hit& h, const Wl& wl) {
    const Frame& f = h.frame;
    if (is_model<MODEL>(m, PT_MAT_LAMBERT)) {
        float c = dot(wi, f.n) * gsign(dot(wo, f.n));
        return gmax(0.0f, c) * INV_PI;
    } else if (is_model<MODEL>(m, PT_MAT_GGX)) {
        f3 wil = to_local(f, wi), wol = to_local(f, wo);
        if (wil.z * wol.z <= 0.0f) return 0.0f;
        float a = m.alpha;
        f3 hh = normalize(wil + wol);
        return mf_G1(wol, a) * mf_D(hh, a) / (4.0f * fabsf(wol.z));
    } else if (is_model<MODEL>(m, PT_MAT_ROUGH_DIELECTRIC)) {   // rough_dielectric.glsl:135-155
        f3 wil = to_local(f, wi), wol = to_local(f, wo);
        if (wol.z <= 0.0f || wil.z == 0.0f) return 0.0f;
        float n_i = ior_of<SPEC>(S, h.region_from, wl);
        float n_t = ior_of<SPEC>(S, h.region_to, wl);
        if (n_i == n_t) return 0.0f;
        float eta = n_i / n_t;
        f3 mh;
        if (!rough_half(wil, wol, eta, mh)) return 0.0f;
        float a = m.alpha;
        float cos_om = gclamp(fabsf(dot(wol, mh)), 1e-6f, 1.0f);
        float F = dielectric_fresnel(cos_om, eta);
        if (wil.z > 0.0f) return F * mf_G1(wol, a) * mf_D(mh, a) / (4.0f * wol.z);
        float denom = dot(wil, mh) + dot(wol, mh) * eta;
        denom *= denom;
        return (1.0f - F) * mf_G1(wol, a) * mf_D(mh, a) * cos_om / wol.z * fabsf(dot(wil, mh)) / denom;
    }
    (void)S; (void)wl;
    return 0.0f;
}
// ---------------------------------------------------------------- environment
// The equirect chart + table sampler (components/env/sampler_cdf.glsl, equirect/*.glsl); the sky
// bodies are pt_sky_body (sky.h).
// spectral_coefficients (lookup.glsl:16-38): the atlas's trilinear lookup for a bounded rgb whose
// largest channel is z; x, y the other two over z (the division before the scale), the z node
// by a six-step search of the 64 nodes.
PT_DEV void atlas_coefficients(const DevScene& S, f3 rgb, float* out) {
    const float v[3] = {rgb.x, rgb.y, rgb.z};
    const int ch = rgb.x >= rgb.y && rgb.x >= rgb.z ? 0 : (rgb.y >= rgb.z ? 1 : 2);
    const float z = v[ch];
    const float x = (v[(ch + 1) % 3] / z) * 63.0f, y = (v[(ch + 2) % 3] / z) * 63.0f;
    int ix = (int)x, iy = (int)y;
    ix = ix < 62 ? ix : 62; iy = iy < 62 ? iy : 62;
    const float fx = x - (float)ix, fy = y - (float)iy;
    int lo = 0, hi = 63;
    for (int step = 0; step < 6; ++step) {
        const int mid = (lo + hi) / 2;
        if (S.spec_znodes[mid] <= z) lo = mid; else hi = mid;
    }
    const float z0 = S.spec_znodes[lo], z1 = S.spec_znodes[lo + 1];
    const float fz = (z - z0) / (z1 - z0);
    const float4* A = S.spec_atlas + (size_t)ch * 64 * 64 * 64;
    float4 n[2][2][2];
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++) n[dz][dy][dx] = A[((size_t)(lo + dz) * 64 + (iy + dy)) * 64 + (ix + dx)];
    for (int c = 0; c < 3; c++) {
        float p[2];
        for (int dz = 0; dz < 2; dz++)
            p[dz] = glsl_mix(glsl_mix(lane(n[dz][0][0], c), lane(n[dz][0][1], c), fx),
                             glsl_mix(lane(n[dz][1][0], c), lane(n[dz][1][1], c), fx), fy);
        out[c] = glsl_mix(p[0], p[1], fz);
    }
}
// color_lift_emission of a runtime RGB (lookup.glsl:50-60): scale = 2 max(rgb); an exact grey
// rgb / scale is Spectrum(value), any other goes through the atlas; times scale, times D65.
PT_DEV sp4 lift_emission_rgb(const DevScene& S, f3 c, const Wl& wl) {
    pt_spectrum r{};
    r.scaled = 1;
    r.d65 = 1;
    r.scale = 2.0f * fmaxf(c.x, fmaxf(c.y, c.z));
    if (r.scale == 0.0f) {
        r.kind = PT_SPEC_ZERO;
    } else {
        const f3 b = mk3(c.x / r.scale, c.y / r.scale, c.z / r.scale);
        if (b.x == b.y && b.y == b.z) {
            r.kind = PT_SPEC_GREY;
            r.value = b.x;
        } else {
            r.kind = PT_SPEC_SIGMOID;
            atlas_coefficients(S, b, r.c);
        }
    }
    return spectrum_eval(S, r, wl);
}
// Environment radiance. The constant sky is color * u_envIntensity in either colour model
// (features/environment.ts:119-133); the procedural sky is its body at the direction turned
// into the table's azimuth (env_rotate_cs), lifted as an emission in a spectral program, times
// u_envIntensity (features/environment.ts:207-213).
template <bool SPEC>
PT_DEV sp4 env_radiance(const DevScene& S, f3 dir, const Wl& wl) {
    if (S.env_type == PT_ENV_CONSTANT)
        return lift<SPEC>(S, &S.env_color.x, S.env_s_color, wl) * S.env_intensity;
    if (S.env_type == PT_ENV_IMAGE) {
        // texture(u_envMap, equirect_uv(dir)) with NEAREST filtering and clamped edges
        const f3 n = normalize(dir);
        const float phi = (n.x == 0.0f && n.z == 0.0f ? 0.0f : ptm_atan2(n.z, n.x)) + S.env_rotation;
        const float theta = ptm_atan2(sqrtf(n.x * n.x + n.z * n.z), n.y);
        const float u = ptm_fract(phi * (1.0f / PTM_TWO_PI) + 0.5f), v = theta * (1.0f / PTM_PI);
        const int W = S.env_img_w, H = S.env_img_h;
        int i = (int)(u * (float)W); i = i < 0 ? 0 : (i > W - 1 ? W - 1 : i);
        int j = (int)(v * (float)H); j = j < 0 ? 0 : (j > H - 1 ? H - 1 : j);
        const float* t = S.env_img + 3 * ((size_t)j * W + i);
        const f3 c = mk3(t[0], t[1], t[2]);
        if (!SPEC) return rgb4(c * S.env_intensity);
        return lift_emission_rgb(S, c, wl) * S.env_intensity;
    }
    if (S.env_type != PT_ENV_SKY) return sgrey(0.0f);
    f3 d = normalize(dir);
    float cx = S.env_rot_c, sy = S.env_rot_s;
    d = mk3(cx * d.x - sy * d.z, d.y, sy * d.x + cx * d.z);
    float c[3];
    pt_sky_body(S.env_body, d.x, d.y, d.z, c);
    if (!SPEC) return rgb4(mk3(c[0], c[1], c[2]) * S.env_intensity);
    return lift_emission_rgb(S, mk3(c[0], c[1], c[2]), wl) * S.env_intensity;
}
PT_DEV float env_texel_dOmega(const DevScene& S, int j, int W, int H) {
    if (S.env_chart == PT_CHART_OCTAHEDRAL) return (4.0f * PTM_PI) / ((float)W * (float)H);   // octahedral.glsl
    float theta = PTM_PI * ((float)j + 0.5f) / (float)H;
    return (PTM_TWO_PI / (float)W) * 2.0f * ptm_sin(theta) * ptm_sin(0.5f * PTM_PI / (float)H);
}
PT_DEV float env_table_pdf(const DevScene& S, f3 dir) {
    const int W = S.env_w, H = S.env_h;
    float u, v;
    if (S.env_chart == PT_CHART_OCTAHEDRAL) ptm_oct_uv(dir.x, dir.y, dir.z, S.env_rot_c, S.env_rot_s, &u, &v);
    else {
        f3 n = normalize(dir);
        float phi = (n.x == 0.0f && n.z == 0.0f ? 0.0f : ptm_atan2(n.z, n.x)) + S.env_rotation;
        float theta = ptm_atan2(sqrtf(n.x * n.x + n.z * n.z), n.y);
        u = ptm_fract(phi * (1.0f / PTM_TWO_PI) + 0.5f);
        v = theta * (1.0f / PTM_PI);
    }
    int i = (int)(u * (float)W); i = i < 0 ? 0 : (i > W - 1 ? W - 1 : i);
    int j = (int)(v * (float)H); j = j < 0 ? 0 : (j > H - 1 ? H - 1 : j);
    float dMarg = S.env_marg[j] - (j > 0 ? S.env_marg[j - 1] : 0.0f);
    float dCond = S.env_cond[(size_t)j * W + i] - (i > 0 ? S.env_cond[(size_t)j * W + i - 1] : 0.0f);
    return (dCond * dMarg) / env_texel_dOmega(S, j, W, H);
}
PT_DEV float env_uniform_fraction(const DevScene& S) {
    if (S.env_compensated) return 0.0f;
    return S.env_marg[S.env_h - 1] > 0.0f ? 0.0625f : 1.0f;
}
PT_DEV float environment_pdf(const DevScene& S, f3 dir) {
    if (S.env_type == PT_ENV_CONSTANT) return 1.0f / (4.0f * PTM_PI);   // CONSTANT_PDF
    float a = env_uniform_fraction(S);
    return (1.0f - a) * env_table_pdf(S, dir) + a / (4.0f * PTM_PI);
}
// The sampler's search (sampler_cdf.glsl): the first index in [0, n - 1] whose CDF value exceeds
// u, n - 1 when none does. The guide narrows it to [g[b], g[b + 1]] for b = floor(u K): the first
// index above u lies between those for u in [b / K, (b + 1) / K), and the search over the narrowed
// range returns it (about one step on average rather than log2 n dependent loads).
PT_DEV int cdf_search(const float* c, int n, float u, const int* g, int k) {
    int lo = 0, hi = n - 1;
    if (u >= 0.0f && u < 1.0f) {
        const int b = (int)(u * (float)k);
        lo = min(g[b], n - 1);
        hi = min(g[b + 1], n - 1);
    }
    while (lo < hi) { const int mid = (lo + hi) >> 1; if (u >= c[mid]) lo = mid + 1; else hi = mid; }
    return lo;
}
PT_DEV f3 env_sample_table(const DevScene& S, float x0, float x1) {
    const int W = S.env_w, H = S.env_h;
    const int j = cdf_search(S.env_marg, H, x0, S.env_guide_marg, S.env_km);
    float mLo = (j > 0) ? S.env_marg[j - 1] : 0.0f;
    float dMarg = S.env_marg[j] - mLo;
    float vFrac = (dMarg > 0.0f) ? gclamp((x0 - mLo) / dMarg, 0.0f, 0.9999999f) : 0.5f;
    const float* row = S.env_cond + (size_t)j * W;
    const int i = cdf_search(row, W, x1, S.env_guide_cond + (size_t)j * (S.env_kc + 1), S.env_kc);
    float cLo = (i > 0) ? row[i - 1] : 0.0f;
    float dCond = row[i] - cLo;
    float uFrac = (dCond > 0.0f) ? gclamp((x1 - cLo) / dCond, 0.0f, 0.9999999f) : 0.5f;
    if (S.env_chart == PT_CHART_OCTAHEDRAL) {         // octahedral.glsl env_chart_sample
        f3 d;
        ptm_oct_dir(((float)i + uFrac) / (float)W, ((float)j + vFrac) / (float)H, S.env_rot_c, S.env_rot_s, &d.x, &d.y, &d.z);
        return d;
    }
    // env_chart_sample (equirect.glsl:36-48)
    float phi = (((float)i + uFrac) / (float)W - 0.5f) * PTM_TWO_PI - S.env_rotation;
    bool north = 2 * j < H;
    float rowf = north ? (float)j : (float)(H - j - 1);
    float a = ptm_sin(0.5f * PTM_PI * rowf / (float)H);
    float b = ptm_sin(0.5f * PTM_PI * (rowf + 1.0f) / (float)H);
    float tt = north ? vFrac : 1.0f - vFrac;
    float hh = a * a * (1.0f - tt) + b * b * tt;
    float y = (north ? 1.0f : -1.0f) * (1.0f - 2.0f * hh);
    float r = 2.0f * sqrtf(hh * (1.0f - hh));
    float sn, cs; ptm_sincos(phi, &sn, &cs);
    return mk3(r * cs, y, r * sn);
}

// ---------------------------------------------------------------- lights
struct LSample { f3 wi; float distance; sp4 radiance; float pdf; int flags; int light_id; };

template <bool SPEC>
PT_DEV LSample light_sample_one(const DevScene& S, const pt_light& l, f3 p, float x0, float x1, const Wl& wl) {
    LSample ls;
    ls.light_id = -1;
    if (l.kind == PT_LIGHT_QUAD) {              // lights/quad/quad.glsl:14-30
        f3 q = (ld3(l.corner) + ld3(l.edge1) * x0) + ld3(l.edge2) * x1;
        f3 d = q - p;
        float d2 = dot(d, d);
        ls.wi = d * (1.0f / sqrtf(d2));
        ls.distance = sqrtf(d2);
        ls.radiance = lift<SPEC>(S, l.radiance, l.s_radiance, wl);
        ls.flags = 0;
        float cos_l = dot(ld3(l.normal), -ls.wi);
        if (cos_l <= 0.0f) { ls.pdf = 0.0f; return ls; }
        ls.pdf = d2 / (l.area * cos_l);
    } else {                                    // lights/point/point.glsl:11-23
        f3 d = ld3(l.corner) - p;
        float d2 = dot(d, d);
        ls.wi = d * (1.0f / sqrtf(d2));
        ls.distance = sqrtf(d2);
        ls.radiance = lift<SPEC>(S, l.radiance, l.s_radiance, wl) / d2;
        ls.pdf = 1.0f; ls.flags = 1;
    }
    return ls;
}
PT_DEV float light_pdf_one(const pt_light& l, f3 p, f3 light_p, f3 wi) {
    if (l.kind != PT_LIGHT_QUAD) return 0.0f;
    float cos_l = dot(ld3(l.normal), -wi);
    if (cos_l <= 0.0f) return 0.0f;
    f3 d = light_p - p;
    return dot(d, d) / (l.area * cos_l);
}
PT_DEV bool env_samplable(const DevScene& S) {
    return S.env_type == PT_ENV_SKY || S.env_type == PT_ENV_IMAGE || (S.env_type == PT_ENV_CONSTANT && S.env_samplable);
}
PT_DEV bool env_select_live(const DevScene& S) { return env_samplable(S) && S.n_lights > 0; }
template <bool SPEC>
PT_DEV LSample lighting_sample_finite(const DevScene& S, f3 p, float x0, float x1, const Wl& wl) {
    const int n = S.n_lights;
    if (n == 1) { LSample ls = light_sample_one<SPEC>(S, S.lights[0], p, x0, x1, wl); ls.light_id = 0; return ls; }
    int id = n - 1; float sel = S.lights[n - 1].select_pdf, lo = S.lights[n - 2].cdf;
    for (int i = 0; i < n - 1; i++) {
        if (x0 < S.lights[i].cdf) { id = i; sel = S.lights[i].select_pdf; lo = i == 0 ? 0.0f : S.lights[i - 1].cdf; break; }
    }
    float xr = (x0 - lo) / sel;
    LSample ls = light_sample_one<SPEC>(S, S.lights[id], p, gclamp(xr, 0.0f, 0.9999999f), x1, wl);
    ls.light_id = id;
    ls.pdf *= sel;
    return ls;
}
template <bool SPEC>
PT_DEV LSample environment_sample(const DevScene& S, float x0, float x1, const Wl& wl) {   // sampler_cdf.glsl:61-80
    LSample ls;
    if (S.env_type == PT_ENV_CONSTANT) {   // the constant sampler (features/environment.ts:74-87)
        const float z = 1.0f - 2.0f * x0;
        const float r = sqrtf(gmax(0.0f, 1.0f - z * z));
        const float phi = PTM_TWO_PI * x1;
        float sn, cs; ptm_sincos(phi, &sn, &cs);
        ls.wi = mk3(r * cs, z, r * sn);
        ls.distance = MAX_DIST;
        ls.radiance = env_radiance<SPEC>(S, ls.wi, wl);
        ls.pdf = 1.0f / (4.0f * PTM_PI);
        ls.flags = 0; ls.light_id = -1;
        return ls;
    }
    float a = env_uniform_fraction(S);
    if (x0 < a) {
        float y = 1.0f - 2.0f * (x0 / a);
        float r = sqrtf(gmax(0.0f, 1.0f - y * y));
        float phi = PTM_TWO_PI * x1;
        float sn, cs; ptm_sincos(phi, &sn, &cs);
        ls.wi = mk3(r * cs, y, r * sn);
    } else {
        ls.wi = env_sample_table(S, (x0 - a) / (1.0f - a), x1);
    }
    ls.distance = MAX_DIST;
    ls.radiance = env_radiance<SPEC>(S, ls.wi, wl);
    ls.pdf = environment_pdf(S, ls.wi);
    ls.flags = 0; ls.light_id = -1;
    return ls;
}
// LightQuery (glsl/core/structs.glsl:83-87; pt.ts:125-140): the receiver point, its shading
// normal (0 at a medium event) and whether its BSDF support is the sphere and NEE evaluates it
// (modelTwoSidedShading, materials/index.ts:56-69: the rough dielectric).
struct LQuery { f3 p, n; bool two_sided; };
// light_tree_importance (accel/light_tree/light_tree.glsl:28-77): power over the clamped squared
// distance; with a normal, the box's exact below-horizon cull and the best-case cosine.
PT_DEV float lt_importance(float4 n0, float4 n1, const LQuery& q) {
    const f3 bmin = xyz(n0), bmax = xyz(n1);
    const f3 c = (bmin + bmax) * 0.5f;
    const f3 e = bmax - bmin;
    const f3 d = c - q.p;
    const float d2raw = dot(d, d);
    const float diag2 = 0.25f * dot(e, e);
    const float base = n0.w / gmax(d2raw, gmax(diag2, 1e-8f));
    if (dot(q.n, q.n) < 0.5f) return base;
    const f3 far_c = mk3(q.n.x < 0.0f ? bmin.x : bmax.x, q.n.y < 0.0f ? bmin.y : bmax.y, q.n.z < 0.0f ? bmin.z : bmax.z);
    float h = dot(q.n, far_c - q.p);
    if (q.two_sided) {
        const f3 near_c = (bmin + bmax) - far_c;
        h = gmax(h, -dot(q.n, near_c - q.p));
    } else if (h <= 0.0f) {
        return 0.0f;
    }
    float cos_c = dot(q.n, d) * (1.0f / sqrtf(gmax(d2raw, 1e-12f)));
    if (q.two_sided) cos_c = fabsf(cos_c);
    const f3 dmx = mk3(gmax(q.p.x - bmin.x, bmax.x - q.p.x), gmax(q.p.y - bmin.y, bmax.y - q.p.y),
                       gmax(q.p.z - bmin.z, bmax.z - q.p.z));
    const float lo = h * (1.0f / sqrtf(gmax(dot(dmx, dmx), 1e-12f)));
    return base * gmin(1.0f, gmax(cos_c, lo));
}
constexpr int LT_MAX_DEPTH = 48;
// light_tree_pick (lighting.ts:786-800): importance-proportional descent, a fresh random number
// per level, the product pmf; -1 on a dead descent (both children culled).
PT_DEV int lt_pick(const DevScene& S, const LQuery& q, float xi, Rng& rng, float& pmf) {
    pmf = 1.0f;
    int ni = 0;
    float4 n1 = S.lt_nodes[1];
    for (int lvl = 0; lvl < LT_MAX_DEPTH; lvl++) {
        if (n1.w < 0.0f) break;
        const int L = ni + 1, R = (int)n1.w;
        const float4 l0 = S.lt_nodes[2 * L], l1 = S.lt_nodes[2 * L + 1];
        const float4 r0 = S.lt_nodes[2 * R], r1 = S.lt_nodes[2 * R + 1];
        const float il = lt_importance(l0, l1, q), ir = lt_importance(r0, r1, q);
        const float isum = il + ir;
        if (isum <= 0.0f) { pmf = 0.0f; return -1; }
        const float p_l = il / isum;
        if (xi < p_l) { pmf *= p_l; ni = L; n1 = l1; }
        else { pmf *= (1.0f - p_l); ni = R; n1 = r1; }
        xi = rng.next();
    }
    return (int)(-n1.w - 1.0f);
}
// light_tree_pmf (lighting.ts:802-822): the light's bit trail replayed with the stored query.
PT_DEV float lt_pmf(const DevScene& S, const LQuery& q, int light) {
    const uint2 trail = S.lt_trails[light];
    uint32_t tw = trail.x;
    float pmf = 1.0f;
    int ni = 0;
    float4 n1 = S.lt_nodes[1];
    for (int lvl = 0; lvl < LT_MAX_DEPTH; lvl++) {
        if (n1.w < 0.0f) break;
        if (lvl == 24) tw = trail.y;
        const int L = ni + 1, R = (int)n1.w;
        const float4 l0 = S.lt_nodes[2 * L], l1 = S.lt_nodes[2 * L + 1];
        const float4 r0 = S.lt_nodes[2 * R], r1 = S.lt_nodes[2 * R + 1];
        const float il = lt_importance(l0, l1, q), ir = lt_importance(r0, r1, q);
        const float isum = il + ir;
        if (isum <= 0.0f) return 0.0f;
        const float p_l = il / isum;
        if ((tw & 1u) == 0u) { pmf *= p_l; ni = L; n1 = l1; }
        else { pmf *= (1.0f - p_l); ni = R; n1 = r1; }
        tw >>= 1u;
    }
    return pmf;
}
// The tree's finite stage (generateLightSamplingBvh, lighting.ts:834-870): pick, the light's sample
// at (random(), x1), its pdf times the selection pmf.
template <bool SPEC>
PT_DEV LSample lighting_sample_tree(const DevScene& S, const LQuery& q, float x0, float x1, Rng& rng, const Wl& wl) {
    float sel;
    const int id = lt_pick(S, q, x0, rng, sel);
    if (id < 0) { LSample ls; ls.pdf = 0.0f; return ls; }
    const float lx = rng.next();
    LSample ls = light_sample_one<SPEC>(S, S.lights[id], q.p, lx, x1, wl);
    ls.light_id = id;
    ls.pdf *= sel;
    return ls;
}
// lighting_sample (features/lighting.ts:546-628, 834-895): env-only, finite-only or two-stage; the
// finite stage by the selection CDF or, with tree, the light tree (rng feeds its descent).
template <bool SPEC>
PT_DEV LSample lighting_sample(const DevScene& S, const LQuery& q, float x0, float x1, const Wl& wl, Rng& rng, bool tree) {
    if (S.n_lights == 0) {
        if (env_samplable(S)) return environment_sample<SPEC>(S, x0, x1, wl);
        LSample ls; ls.pdf = 0.0f; return ls;
    }
    if (!env_select_live(S))
        return tree ? lighting_sample_tree<SPEC>(S, q, x0, x1, rng, wl) : lighting_sample_finite<SPEC>(S, q.p, x0, x1, wl);
    float ps = S.env_select_prob;
    if (x0 < ps) {
        LSample ls = environment_sample<SPEC>(S, gclamp(x0 / ps, 0.0f, 0.9999999f), x1, wl);
        ls.pdf *= ps;
        return ls;
    }
    float fx = gclamp((x0 - ps) / (1.0f - ps), 0.0f, 0.9999999f);
    LSample ls = tree ? lighting_sample_tree<SPEC>(S, q, fx, x1, rng, wl) : lighting_sample_finite<SPEC>(S, q.p, fx, x1, wl);
    ls.pdf *= (1.0f - ps);
    return ls;
}
PT_DEV int light_of(const DevScene& S, int region) {
    if (region < 0) return -1;
    for (int i = 0; i < S.n_lights; i++) if (S.lights[i].region == region) return i;
    return -1;
}
// lighting_pdf (lighting.ts:636-673; the tree's, :909-933, replays the light's trail)
PT_DEV float lighting_pdf(const DevScene& S, const LQuery& q, f3 wi, int lid, f3 light_p, bool tree) {
    const pt_light& l = S.lights[lid];
    float sel = tree ? lt_pmf(S, q, lid) : (S.n_lights == 1 ? 1.0f : l.select_pdf);
    if (env_select_live(S)) sel = sel * (1.0f - S.env_select_prob);
    return sel * light_pdf_one(l, q.p, light_p, wi);
}
// ---------------------------------------------------------------- media
// material_of for regions (the ambient region -1 maps to the ambientMedium material),
// the analytic channel-MIS arm (transport/volume/analytic/analytic.glsl:9-51) and HG
// (volume_scattering/hg/hg.glsl:11-42).
PT_DEV int material_of(const DevScene& S, int region) { return region < 0 ? S.ambient_material : S.objects[region].material; }
PT_DEV bool is_null_interface(const DevScene& S, int mat) { return mat >= 0 && S.materials[mat].model == PT_MAT_NULL; }
PT_DEV bool material_has_medium(const DevScene& S, int mat) { return mat >= 0 && S.materials[mat].has_medium; }
PT_DEV sp4 spectrum_exp(sp4 s) { return mk4(ptm_exp(s.x), ptm_exp(s.y), ptm_exp(s.z), ptm_exp(s.w)); }
template <bool SPEC> PT_DEV sp4 sigma_s_of(const DevScene& S, const pt_material& m, const Wl& wl) { return lift<SPEC>(S, m.sigma_s, m.s_sigma_s, wl); }
template <bool SPEC> PT_DEV sp4 sigma_a_of(const DevScene& S, const pt_material& m, const Wl& wl) { return lift<SPEC>(S, m.sigma_a, m.s_sigma_a, wl); }
struct MSample { bool scattered; float t; sp4 weight; sp4 radiance; };
template <bool SPEC>
PT_DEV MSample medium_sample_analytic(const DevScene& S, const pt_material& m, float t_max, float x0, float x1, const Wl& wl) {
    MSample ms;
    sp4 ss = sigma_s_of<SPEC>(S, m, wl);
    sp4 st = sigma_a_of<SPEC>(S, m, wl) + ss;
    constexpr int N = SPEC ? 4 : 3;
    int c = (int)(x0 * (float)N); if (c > N - 1) c = N - 1;     // spectrum_sample_channel
    float sc = spectrum_comp(st, c);
    float t = sc == 0.0f ? t_max : -ptm_log(1.0f - x1) / sc;
    if (t < t_max) {
        sp4 tr = spectrum_exp((-st) * t);
        float pdf = spectrum_average<SPEC>(st * tr);
        ms.scattered = true; ms.t = t;
        ms.weight = (ss * tr) / gmax(pdf, 1e-20f);
    } else {
        sp4 tr = spectrum_exp((-st) * t_max);
        float pdf = spectrum_average<SPEC>(tr);
        ms.scattered = false; ms.t = t_max;
        ms.weight = tr / gmax(pdf, 1e-20f);
    }
    ms.radiance = sgrey(0.0f);
    return ms;
}
template <bool SPEC> PT_DEV sp4 medium_emission_of(const DevScene& S, const pt_material& m, const Wl& wl) {
    return lift<SPEC>(S, m.medium_emission, m.s_medium_emission, wl);
}
PT_DEV bool medium_emits(const pt_material& m) {
    return m.medium_emission[0] != 0.0f || m.medium_emission[1] != 0.0f || m.medium_emission[2] != 0.0f;
}
// integral_0^len exp(-sigma s) ds, Taylor form below optical depth 0.1 (analytic/emission.glsl:4-13)
PT_DEV float absorption_integral1(float sigma, float len) {
    const float tau = sigma * len;
    return tau < 0.1f ? len * (1.0f + tau * (-0.5f + tau * (1.0f / 6.0f + tau * (-1.0f / 24.0f + tau / 120.0f))))
                      : (1.0f - ptm_exp(-tau)) / sigma;
}
// Deterministic absorbing arm (features/materials.ts:708-718): no collision, weight
// exp(-sigma_a t_max), emission eps * integral of the attenuation along the segment.
template <bool SPEC>
PT_DEV MSample medium_sample_absorb(const DevScene& S, const pt_material& m, float t_max, const Wl& wl) {
    MSample ms;
    const sp4 sa = sigma_a_of<SPEC>(S, m, wl);
    ms.scattered = false; ms.t = t_max;
    ms.weight = spectrum_exp((-sa) * t_max);
    ms.radiance = sgrey(0.0f);
    if (medium_emits(m))
        ms.radiance = medium_emission_of<SPEC>(S, m, wl) *
                      mk4(absorption_integral1(sa.x, t_max), absorption_integral1(sa.y, t_max),
                          absorption_integral1(sa.z, t_max), absorption_integral1(sa.w, t_max));
    return ms;
}
// Delta tracking on a constant medium at the scalar majorant (delta_tracking.glsl:46-120: Kutz
// et al. 2017 Alg. 4, history-aware average probabilities; tracking_cap bounds the walk):
// emission per tentative collision before the lottery, absorption terminates, null collisions
// reweight and draw again from the path's stream.
template <bool SPEC>
PT_DEV MSample medium_sample_delta(const DevScene& S, const pt_material& m, float t_max, float x0, float x1, Rng& rng,
                                   const Wl& wl) {
    MSample ms;
    ms.scattered = false; ms.t = t_max; ms.weight = sgrey(1.0f); ms.radiance = sgrey(0.0f);
    const float sb = SPEC ? m.majorant[1] : m.majorant[0];
    const sp4 sa = sigma_a_of<SPEC>(S, m, wl), ss = sigma_s_of<SPEC>(S, m, wl), eps = medium_emission_of<SPEC>(S, m, wl);
    const sp4 st = sa + ss;
    const sp4 sn = mk4(gmax(sb - st.x, 0.0f), gmax(sb - st.y, 0.0f), gmax(sb - st.z, 0.0f), gmax(sb - st.w, 0.0f));
    sp4 w = sgrey(1.0f);
    float t = 0.0f, xd = x0, xe = x1;
    const float lam = sb * t_max;
    const int cap = (int)ceilf(lam + 6.5f * sqrtf(lam) + 12.0f);
    for (int i = 0; i < cap; i++) {
        t += -ptm_log(1.0f - xd) / sb;
        if (t >= t_max) { ms.weight = w; return ms; }
        ms.radiance = ms.radiance + (w * eps) / sb;
        float pa = spectrum_average<SPEC>(sa * w), ps = spectrum_average<SPEC>(ss * w), pn = spectrum_average<SPEC>(sn * w);
        const float c = pa + ps + pn;
        if (c <= 0.0f) { ms.weight = sgrey(0.0f); return ms; }
        pa /= c; ps /= c; pn /= c;
        if (xe < pa) { ms.weight = sgrey(0.0f); return ms; }
        if (xe < pa + ps) {
            ms.scattered = true; ms.t = t;
            ms.weight = (w * ss) / (sb * gmax(ps, 1e-20f));
            return ms;
        }
        w = w * (sn / (sb * gmax(pn, 1e-20f)));
        xd = rng.next(); xe = rng.next();
    }
    ms.weight = w;
    return ms;
}
template <bool SPEC>
PT_DEV sp4 medium_transmittance(const DevScene& S, const pt_material& m, float len, const Wl& wl) {
    return spectrum_exp((-(sigma_a_of<SPEC>(S, m, wl) + sigma_s_of<SPEC>(S, m, wl))) * len);
}
PT_DEV float hg_value(f3 wi, f3 wo, float g) {
    float c = dot(wi, -wo);
    float d = 1.0f + g * g - 2.0f * g * c;
    return (1.0f - g * g) / (4.0f * PTM_PI * d * sqrtf(gmax(d, 1e-8f)));
}
// rayleigh_eval (volume_scattering/rayleigh/rayleigh.glsl): (3 / 16 pi)(1 + c^2), c measured from
// the propagation direction; 3 / (16 PI) folded in fp32.
PT_DEV float rayleigh_value(f3 wi, f3 wo) {
    const float c = dot(wi, -wo);
    return (3.0f / (16.0f * PTM_PI)) * (1.0f + c * c);
}
// The medium's phase function (kernel_phase.glsl dispatch): value and pdf (grayscale: equal).
PT_DEV float phase_value(int model, f3 wi, f3 wo, float g) {
    return model == PT_PHASE_RAYLEIGH ? rayleigh_value(wi, wo) : hg_value(wi, wo, g);
}
template <bool SPEC>
PT_DEV float phase_pdf(int model, f3 wi, f3 wo, float g) { return spectrum_average<SPEC>(sgrey(phase_value(model, wi, wo, g))); }
// hg_sample (hg.glsl) and rayleigh_sample (rayleigh.glsl: the marginal (3/8)(1 + mu^2) inverted in
// closed form, Cardano's root of mu^3 + 3 mu + (4 - 8 xi) = 0), then the same frame around the
// propagation direction; the weight is 1 for both.
template <bool SPEC>
PT_DEV f3 phase_sample(int model, f3 wo, float g, float x0, float x1, float& pdf) {
    float cos_theta;
    if (model == PT_PHASE_RAYLEIGH) {
        const float u = 4.0f * x0 - 2.0f;
        const float D = sqrtf(u * u + 1.0f);
        cos_theta = ptm_pow(u + D, 1.0f / 3.0f) - ptm_pow(D - u, 1.0f / 3.0f);
        cos_theta = gclamp(cos_theta, -1.0f, 1.0f);
    } else if (fabsf(g) < 1e-3f) {
        cos_theta = 1.0f - 2.0f * x0;
    } else {
        float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * x0);
        cos_theta = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    float sin_theta = sqrtf(gmax(0.0f, 1.0f - cos_theta * cos_theta));
    float phi = PTM_TWO_PI * x1;
    f3 nw = -wo;
    // build_basis (glsl/core/math.glsl:63-67)
    f3 up = fabsf(nw.y) < 0.999f ? mk3(0.0f, 1.0f, 0.0f) : mk3(1.0f, 0.0f, 0.0f);
    f3 t = normalize(cross(up, nw));
    f3 b = cross(nw, t);
    float sn, cs; ptm_sincos(phi, &sn, &cs);
    f3 wi = normalize((t * (sin_theta * cs) + b * (sin_theta * sn)) + nw * cos_theta);
    pdf = phase_pdf<SPEC>(model, wi, wo, g);
    return wi;
}

PT_DEV float power_heuristic(float pf, float pg) {   // transport/math_mis.glsl:5-13
    float f2 = pf * pf;
    if (f2 >= 1.0e38f) return 1.0f;
    return f2 / gmax(1e-20f, f2 + pg * pg);
}

}  // namespace ptmi
