// lfpm_calib.cpp — calibration ingest (host only): the intrinsics.xml and CalibInfo.xml readers
// of DepthSolver::ReadIntrinsicsFromXML / ReadCalibrationParamsFromXML (DepthSolver.cpp:872-1110,
// 1112-1251) over a small XML reader of this file, and the FixedPosition neighbour table of a
// micro-lens table read that way.  Numbers are parsed with the reference's sscanf formats.
#include <algorithm>
#include <cmath>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lfpm.h"
#include "lfpm_internal.h"

namespace {

struct XNode {
    std::string name, text;  // text: the first character data before any child element
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XNode> kids;
    const char* attr(const char* k) const {
        for (const auto& a : attrs)
            if (a.first == k) return a.second.c_str();
        return nullptr;
    }
};

// Element tree of an XML document: prolog, comments, DOCTYPE and processing instructions are
// skipped; the five predefined entities and numeric character references are decoded.
struct XParser {
    const std::string& s;
    size_t i = 0;
    std::string err;
    explicit XParser(const std::string& src) : s(src) {}

    bool starts(const char* t) const { return s.compare(i, std::strlen(t), t) == 0; }
    void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
    bool skip_to(const char* t) {
        const size_t k = s.find(t, i);
        if (k == std::string::npos) return false;
        i = k + std::strlen(t);
        return true;
    }
    bool misc() {  // comments, PIs, DOCTYPE
        for (;;) {
            ws();
            if (starts("<!--")) { if (!skip_to("-->")) return false; }
            else if (starts("<?")) { if (!skip_to("?>")) return false; }
            else if (starts("<!")) { if (!skip_to(">")) return false; }
            else return true;
        }
    }
    static void decode(std::string& t) {
        std::string o;
        for (size_t k = 0; k < t.size(); ++k) {
            if (t[k] != '&') { o += t[k]; continue; }
            const size_t e = t.find(';', k);
            if (e == std::string::npos) { o += t[k]; continue; }
            const std::string ent = t.substr(k + 1, e - k - 1);
            if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "amp") o += '&';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else if (!ent.empty() && ent[0] == '#') {
                const long c = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X')
                                   ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                   : std::strtol(ent.c_str() + 1, nullptr, 10);
                if (c > 0 && c < 128) o += (char)c;
            } else {
                o += t.substr(k, e - k + 1);
            }
            k = e;
        }
        t.swap(o);
    }
    std::string name() {
        const size_t b = i;
        while (i < s.size() && !std::isspace((unsigned char)s[i]) && s[i] != '>' && s[i] != '/' &&
               s[i] != '=')
            ++i;
        return s.substr(b, i - b);
    }
    bool element(XNode& n, int depth) {
        if (depth > 64) { err = "nesting too deep"; return false; }
        if (!starts("<")) { err = "expected an element"; return false; }
        ++i;
        n.name = name();
        if (n.name.empty()) { err = "empty element name"; return false; }
        for (;;) {
            ws();
            if (i >= s.size()) { err = "unterminated tag"; return false; }
            if (starts("/>")) { i += 2; return true; }
            if (s[i] == '>') { ++i; break; }
            std::string k = name();
            ws();
            if (i >= s.size() || s[i] != '=') { err = "attribute without value"; return false; }
            ++i;
            ws();
            if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) { err = "unquoted attribute"; return false; }
            const char q = s[i++];
            const size_t e = s.find(q, i);
            if (e == std::string::npos) { err = "unterminated attribute"; return false; }
            std::string v = s.substr(i, e - i);
            decode(v);
            n.attrs.emplace_back(k, v);
            i = e + 1;
        }
        bool have_text = false;
        for (;;) {
            if (i >= s.size()) { err = "unterminated element " + n.name; return false; }
            if (starts("<!--")) { if (!skip_to("-->")) { err = "unterminated comment"; return false; } continue; }
            if (starts("<![CDATA[")) {
                const size_t b = i + 9, e = s.find("]]>", b);
                if (e == std::string::npos) { err = "unterminated CDATA"; return false; }
                if (!have_text && n.kids.empty()) { n.text += s.substr(b, e - b); }
                i = e + 3;
                continue;
            }
            if (starts("<?")) { if (!skip_to("?>")) { err = "unterminated PI"; return false; } continue; }
            if (starts("</")) {
                i += 2;
                const std::string close = name();
                ws();
                if (close != n.name || i >= s.size() || s[i] != '>') {
                    err = "mismatched closing tag " + close + " for " + n.name;
                    return false;
                }
                ++i;
                return true;
            }
            if (s[i] == '<') {
                n.kids.emplace_back();
                if (!element(n.kids.back(), depth + 1)) return false;
                have_text = have_text || !n.text.empty();
                continue;
            }
            const size_t e = s.find('<', i);
            std::string t = s.substr(i, (e == std::string::npos ? s.size() : e) - i);
            i = e == std::string::npos ? s.size() : e;
            if (n.kids.empty() && !have_text) {
                decode(t);
                n.text += t;
            }
        }
    }
    bool document(XNode& root) {
        if (!misc()) { err = "unterminated prolog"; return false; }
        return element(root, 0);
    }
};

int load_xml(const char* path, XNode& root) {
    if (!path) return lfpm::set_error(LFPM_E_INVALID, "xml: null path");
    FILE* f = std::fopen(path, "rb");
    if (!f) return lfpm::set_error(LFPM_E_INVALID, "xml: cannot open %s", path);
    std::string src;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) src.append(buf, n);
    std::fclose(f);
    XParser p(src);
    if (!p.document(root))
        return lfpm::set_error(LFPM_E_INVALID, "xml: %s: %s", path, p.err.c_str());
    return LFPM_OK;
}

const XNode* first_child(const XNode& n) { return n.kids.empty() ? nullptr : &n.kids[0]; }

// the three element texts under the first child of n (a "_" group of values), as %lf
void three(const XNode& n, double* out, int count) {
    const XNode* g = first_child(n);
    if (!g) return;
    for (int k = 0; k < count && k < (int)g->kids.size(); ++k)
        std::sscanf(g->kids[k].text.c_str(), "%lf", &out[k]);
}

}  // namespace

extern "C" {

int lfpm_read_intrinsics_xml(const char* path, lfpm_intrinsics* out) {
    if (!out) return lfpm::set_error(LFPM_E_INVALID, "intrinsics: null output");
    XNode root;
    if (int rc = load_xml(path, root)) return rc;
    if (root.name != "_")
        return lfpm::set_error(LFPM_E_INVALID, "intrinsics: root element %s, expected _",
                               root.name.c_str());
    std::memset(out, 0, sizeof *out);
    for (const XNode& g : root.kids) {
        if (g.name == "D") {
            std::sscanf(g.text.c_str(), "%f", &out->D);
        } else if (g.name == "d") {
            std::sscanf(g.text.c_str(), "%f", &out->d);
        } else if (g.name == "sensor") {
            for (const XNode& e : g.kids) {
                if (e.name == "pose") {
                    for (const XNode& c : e.kids) {
                        if (c.name == "rotation") {
                            // three "_" columns of three values: column j -> R[r][j]
                            for (int j = 0; j < 3 && j < (int)c.kids.size(); ++j) {
                                const XNode& col = c.kids[j];
                                if (col.name != "_") continue;
                                for (int r = 0; r < 3 && r < (int)col.kids.size(); ++r)
                                    std::sscanf(col.kids[r].text.c_str(), "%lf",
                                                &out->sensor_rotation[r * 3 + j]);
                            }
                        } else if (c.name == "translation") {
                            const XNode* t = first_child(c);
                            if (t && t->name == "_")
                                for (int k = 0; k < 3 && k < (int)t->kids.size(); ++k)
                                    std::sscanf(t->kids[k].text.c_str(), "%lf",
                                                &out->sensor_translation[k]);
                        }
                    }
                } else if (e.name == "scale") {
                    std::sscanf(e.text.c_str(), "%f", &out->sensor_pixel_size);
                }
            }
        } else if (g.name == "main_lens") {
            for (const XNode& e : g.kids)
                if (e.name == "f") std::sscanf(e.text.c_str(), "%f", &out->main_lens_f);
        } else if (g.name == "mia") {
            for (const XNode& e : g.kids) {
                if (e.name != "mesh") continue;
                for (const XNode& c : e.kids) {
                    if (c.name == "height") {
                        std::sscanf(c.text.c_str(), "%d", &out->mla_v_size);
                    } else if (c.name == "width") {
                        std::sscanf(c.text.c_str(), "%d", &out->mla_u_size);
                    } else if (c.name == "pitch") {
                        // every grandchild under every "_" child is read in turn: the last wins
                        for (const XNode& cc : c.kids)
                            if (cc.name == "_")
                                for (const XNode& gc : cc.kids)
                                    std::sscanf(gc.text.c_str(), "%f", &out->baseline);
                    }
                }
            }
        } else if (g.name == "distortions") {
            for (const XNode& e : g.kids) {
                if (e.name == "depth") three(e, out->distortion_depth, 3);
                else if (e.name == "radial") three(e, out->distortion_radial, 3);
                else if (e.name == "tangential") three(e, out->distortion_tangential, 2);
            }
        }
    }
    return LFPM_OK;
}

int lfpm_read_calib_xml(const char* path, int capacity, int* tilekeys, float* centers,
                        int* lens_type, float* areas, float* blur_radius, int* count) {
    if (!count) return lfpm::set_error(LFPM_E_INVALID, "calib: null count");
    XNode root;
    if (int rc = load_xml(path, root)) return rc;
    if (root.name != "Calib")
        return lfpm::set_error(LFPM_E_INVALID, "calib: root element %s, expected Calib",
                               root.name.c_str());
    struct Mla { int row, col, type; float cx, cy, area; };
    std::vector<Mla> mla;
    float radius[3] = {0.0f, 0.0f, 0.0f};  // std::map<int, float>::operator[] default
    for (const XNode& g : root.kids) {
        if (g.name == "BlurInfo") {
            for (const XNode& e : g.kids) {
                if (e.name != "MLAType") continue;
                float r = -1.0f;
                std::sscanf(e.text.c_str(), "%f", &r);
                int type = -1;
                if (const char* id = e.attr("id")) std::sscanf(id, "%i", &type);
                if (type >= 0 && type < 3) radius[type] = r;
            }
        } else if (g.name == "MLAInfo") {
            for (const XNode& e : g.kids) {
                if (e.name != "MLA") continue;
                Mla m{-1, -1, -1, 0.0f, 0.0f, 0.0f};
                for (const XNode& c : e.kids) {
                    if (c.name == "key") std::sscanf(c.text.c_str(), "%i,%i", &m.row, &m.col);
                    else if (c.name == "center")
                        std::sscanf(c.text.c_str(), "%f,%f", &m.cx, &m.cy);
                    else if (c.name == "MLAType") std::sscanf(c.text.c_str(), "%i", &m.type);
                    else if (c.name == "aera") std::sscanf(c.text.c_str(), "%f", &m.area);
                }
                mla.push_back(m);
            }
        }
    }
    // the MLA map is keyed by (row, col) (QuadTreeTileKey order, Common/QuadTree.h:137-148,
    // compared by value through QuadTreeTileKeyMapCmpLess); a repeated key keeps its last entry,
    // as the reference's m_MLA_info_map[ptrKey] = ptrInfo assignment does (DepthSolver.cpp:1242):
    // stable sort, then the last of each run of equal keys
    std::stable_sort(mla.begin(), mla.end(), [](const Mla& a, const Mla& b) {
        return a.row != b.row ? a.row < b.row : a.col < b.col;
    });
    {
        std::vector<Mla> last;
        last.reserve(mla.size());
        for (size_t i = 0; i < mla.size(); ++i)
            if (i + 1 == mla.size() || mla[i + 1].row != mla[i].row || mla[i + 1].col != mla[i].col)
                last.push_back(mla[i]);
        mla.swap(last);
    }
    const int n = (int)mla.size();
    *count = n;
    if (blur_radius)
        for (int t = 0; t < 3; ++t) blur_radius[t] = radius[t];
    if (!tilekeys && !centers && !lens_type && !areas) return LFPM_OK;
    if (capacity < n)
        return lfpm::set_error(LFPM_E_INVALID, "calib: capacity %d < %d lenses", capacity, n);
    for (int v = 0; v < n; ++v) {
        if (tilekeys) { tilekeys[2 * v] = mla[v].col; tilekeys[2 * v + 1] = mla[v].row; }
        if (centers) { centers[2 * v] = mla[v].cx; centers[2 * v + 1] = mla[v].cy; }
        if (lens_type) lens_type[v] = mla[v].type;
        if (areas) areas[v] = mla[v].area;
    }
    return LFPM_OK;
}

int lfpm_table_fixed_neighbors(const int* tilekeys, const float* centers, int num_mis,
                               int lattice_cols, int lattice_rows, int max_neighbors,
                               int* neighbor_ids, int* neighbor_counts, int* garbage) {
    if (!tilekeys || !centers || num_mis < 0 || lattice_cols <= 0 || lattice_rows <= 0)
        return lfpm::set_error(LFPM_E_INVALID, "table neighbours: bad arguments");
    std::vector<int> cell((size_t)lattice_cols * lattice_rows, -1), mi_i(num_mis), mi_j(num_mis);
    std::vector<float> cx(num_mis), cy(num_mis);
    for (int v = 0; v < num_mis; ++v) {
        const int X = tilekeys[2 * v], Y = tilekeys[2 * v + 1];
        if (X < 0 || Y < 0 || X >= lattice_cols || Y >= lattice_rows)
            return lfpm::set_error(LFPM_E_INVALID, "table neighbours: key (%d, %d) outside the "
                                   "%d x %d lattice", X, Y, lattice_cols, lattice_rows);
        int& c = cell[(size_t)Y * lattice_cols + X];
        if (c >= 0)
            return lfpm::set_error(LFPM_E_INVALID, "table neighbours: key (%d, %d) repeated", X, Y);
        c = v;
        mi_i[v] = X;
        mi_j[v] = Y;
        cx[v] = centers[2 * v];
        cy[v] = centers[2 * v + 1];
    }
    return lfpm::fixed_neighbors_core(num_mis, mi_i.data(), mi_j.data(), cx.data(), cy.data(),
                                      cell.data(), lattice_cols, lattice_rows, max_neighbors,
                                      neighbor_ids, neighbor_counts, garbage);
}

int lfpm_table_ring_neighbors(const int* tilekeys, const float* centers, int num_mis,
                              int lattice_cols, int lattice_rows, int rings, int max_neighbors,
                              int* neighbor_ids, int* neighbor_counts) {
    if (!tilekeys || !centers || num_mis < 0 || lattice_cols <= 0 || lattice_rows <= 0 ||
        rings < 1 || rings > 5 || max_neighbors < 1 || max_neighbors > LFPM_MAX_NEIGHBORS)
        return lfpm::set_error(LFPM_E_INVALID, "ring neighbours: bad arguments");
    std::vector<int> cell((size_t)lattice_cols * lattice_rows, -1);
    for (int v = 0; v < num_mis; ++v) {
        const int X = tilekeys[2 * v], Y = tilekeys[2 * v + 1];
        if (X < 0 || Y < 0 || X >= lattice_cols || Y >= lattice_rows)
            return lfpm::set_error(LFPM_E_INVALID, "ring neighbours: key (%d, %d) outside the "
                                   "%d x %d lattice", X, Y, lattice_cols, lattice_rows);
        int& c = cell[(size_t)Y * lattice_cols + X];
        if (c >= 0)
            return lfpm::set_error(LFPM_E_INVALID, "ring neighbours: key (%d, %d) repeated", X, Y);
        c = v;
    }
    // ring walk (SelectNeighbors::CollectMIANeighImagesForMatch, MINeighbor/SelectNeighbors.cpp:
    // 113-193, circles 1..rings of the hex lattice): keys at cube distance 1..rings in the
    // odd-r offset lattice (odd key rows sit half a pitch to the right); the score of
    // ComputeNeighborScoresForMatch (:770-830, weights a = 1, b = c = 0) is 1 - the normalised
    // baseline, so SortNeighScoreForMatch (MVStereo/LFDepthInfo.cpp:214-241) orders them by
    // baseline; equal baselines by angle in [0, 2 pi), then collection order.
    struct Cand { int u; float base; double ang; int order; };
    std::vector<Cand> cand;
    for (int v = 0; v < num_mis; ++v) {
        const int X = tilekeys[2 * v], Y = tilekeys[2 * v + 1];
        const int q = X - (Y - (Y & 1)) / 2;
        cand.clear();
        for (int dr = -rings; dr <= rings; ++dr)
            for (int dq = -rings; dq <= rings; ++dq) {
                const int ring = std::max(std::abs(dq), std::max(std::abs(dr), std::abs(dq + dr)));
                if (ring == 0 || ring > rings) continue;
                const int r2 = Y + dr, q2 = q + dq;
                const int X2 = q2 + (r2 - (r2 & 1)) / 2;
                if (X2 < 0 || r2 < 0 || X2 >= lattice_cols || r2 >= lattice_rows) continue;
                const int u = cell[(size_t)r2 * lattice_cols + X2];
                if (u < 0) continue;
                const float dx = centers[2 * u] - centers[2 * v];
                const float dy = centers[2 * u + 1] - centers[2 * v + 1];
                // the reference's float baseline sqrt(pow(dx, 2) + pow(dy, 2)) (pow in double)
                const float b = (float)std::sqrt((double)dx * dx + (double)dy * dy);
                double a = std::atan2((double)dy, (double)dx);
                if (a < 0) a += 2.0 * M_PI;
                cand.push_back(Cand{u, b, a, (int)cand.size()});
            }
        std::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) {
            if (a.base != b.base) return a.base < b.base;
            if (a.ang != b.ang) return a.ang < b.ang;
            return a.order < b.order;
        });
        const int n = std::min((int)cand.size(), max_neighbors);
        for (int k = 0; k < max_neighbors; ++k)
            if (neighbor_ids) neighbor_ids[(size_t)v * max_neighbors + k] = k < n ? cand[k].u : -1;
        if (neighbor_counts) neighbor_counts[v] = n;
    }
    return LFPM_OK;
}

int lfpm_table_origins(const float* centers, int num_mis, int crop, int mla_u_size,
                       int mla_v_size, float baseline, int raw_w, int raw_h, int* origins,
                       int* kept) {
    if (!centers || num_mis < 0 || crop < 1 || !origins || !kept)
        return lfpm::set_error(LFPM_E_INVALID, "table origins: bad arguments");
    // DepthSolver::CreateProblemsAndDisNormals_Frame (DepthSolver.cpp:1894-1905): corners in
    // FP32 from the double offset (crop - 1) / 2; a lens is abandoned when its crop leaves the
    // lens area mla_u_size * baseline x mla_v_size * baseline; the slicing rectangle takes the
    // corner truncated to int (:1977)
    const double half = (crop - 1) * 0.5;
    const float lim_x = (float)mla_u_size * baseline, lim_y = (float)mla_v_size * baseline;
    for (int v = 0; v < num_mis; ++v) {
        const float xl = (float)((double)centers[2 * v] - half);
        const float yl = (float)((double)centers[2 * v + 1] - half);
        const float xr = (float)((double)centers[2 * v] + half);
        const float yr = (float)((double)centers[2 * v + 1] + half);
        bool ok = !(xl < 0 || yl < 0 || xr > lim_x || yr > lim_y);
        const int ox = ok ? (int)xl : 0, oy = ok ? (int)yl : 0;
        // and the crop must lie in the raw frame (the reference would cut past its edge)
        ok = ok && ox + crop <= raw_w && oy + crop <= raw_h;
        origins[2 * v] = ok ? ox : -1;
        origins[2 * v + 1] = ok ? oy : -1;
        kept[v] = ok ? 1 : 0;
    }
    return LFPM_OK;
}

}  // extern "C"
