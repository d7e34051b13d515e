// ingest.cpp -- the reference's input side without GDAL or tinyxml2:
//   * Sentinel-1 annotation XML -> subswath metadata (S1_Initialize,
//     Func.cpp:135-574) with the per-burst Doppler centroid / FM-rate
//     polynomials of ComputeDerampPara (Func.cpp:1868-2031);
//   * precise-orbit EOF -> state vectors (S1_OrbitInitialize, Func.cpp:704-819);
//   * stripped uncompressed TIFF / BigTIFF read and write (TiffRead /
//     TiffWrite, DataTypes.cpp:417-620), and a read that lands in device
//     memory through pinned staging buffers.
// All of it is host code.  A small XML reader replaces tinyxml2: elements,
// attributes, text, comments, CDATA, processing instructions and the five
// predefined entities -- what the annotation and EOF files use.
#include "tops_internal.hpp"

#include <errno.h>
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <string>
#include <vector>

namespace tops {
namespace {

// ---------------------------------------------------------------------------
// XML
// ---------------------------------------------------------------------------
struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::string text;          // first non-blank text run, entities decoded
    bool has_text = false;
    int parent = -1;
    std::vector<int> kids;     // child elements in document order
    int index_in_parent = 0;
};

struct XmlDoc {
    std::vector<XmlNode> nodes;   // nodes[0] is the root element
    std::string error;

    int child(int n, const char* name) const
    {
        if (n < 0) return -1;
        for (int k : nodes[n].kids)
            if (!name || nodes[k].name == name) return k;
        return -1;
    }
    int next_sibling(int n) const
    {
        if (n < 0 || nodes[n].parent < 0) return -1;
        const XmlNode& p = nodes[nodes[n].parent];
        const size_t i = (size_t)nodes[n].index_in_parent + 1;
        return i < p.kids.size() ? p.kids[i] : -1;
    }
    const char* text(int n) const { return n >= 0 && nodes[n].has_text ? nodes[n].text.c_str() : nullptr; }
    const char* first_attr(int n) const
    {
        return n >= 0 && !nodes[n].attrs.empty() ? nodes[n].attrs[0].second.c_str() : nullptr;
    }
};

std::string decode_entities(const char* b, const char* e)
{
    std::string out;
    out.reserve(e - b);
    for (const char* p = b; p < e; p++) {
        if (*p != '&') {
            out += *p;
            continue;
        }
        static const struct { const char* s; char c; } ents[] = {
            {"&lt;", '<'}, {"&gt;", '>'}, {"&amp;", '&'}, {"&quot;", '"'}, {"&apos;", '\''}};
        bool done = false;
        for (auto& en : ents) {
            const size_t n = strlen(en.s);
            if ((size_t)(e - p) >= n && !strncmp(p, en.s, n)) {
                out += en.c;
                p += n - 1;
                done = true;
                break;
            }
        }
        if (!done && p + 2 < e && p[1] == '#') {
            const char* q = p + 2;
            const bool hex = *q == 'x';
            if (hex) q++;
            char* end = nullptr;
            const long v = strtol(q, &end, hex ? 16 : 10);
            if (end && end < e && *end == ';' && v > 0 && v < 128) {
                out += (char)v;
                p = end;
                done = true;
            }
        }
        if (!done) out += '&';
    }
    return out;
}

bool is_blank(const char* b, const char* e)
{
    for (const char* p = b; p < e; p++)
        if (*p != ' ' && *p != '\t' && *p != '\r' && *p != '\n') return false;
    return true;
}

bool parse_xml(const std::string& src, XmlDoc& doc)
{
    const char* p = src.c_str();
    const char* end = p + src.size();
    if (end - p >= 3 && (unsigned char)p[0] == 0xEF && (unsigned char)p[1] == 0xBB &&
        (unsigned char)p[2] == 0xBF)
        p += 3;
    std::vector<int> stack;
    auto fail = [&](const char* what) {
        char b[160];
        snprintf(b, sizeof(b), "%s at byte %ld", what, (long)(p - src.c_str()));
        doc.error = b;
        return false;
    };
    auto add_text = [&](const char* b, const char* e, bool raw) {
        if (stack.empty() || b >= e) return;
        if (!raw && is_blank(b, e)) return;
        XmlNode& n = doc.nodes[stack.back()];
        if (n.has_text) return;   // the first text run (tinyxml2 GetText)
        n.text = raw ? std::string(b, e) : decode_entities(b, e);
        n.has_text = true;
    };
    while (p < end) {
        if (*p != '<') {
            const char* q = (const char*)memchr(p, '<', end - p);
            if (!q) q = end;
            add_text(p, q, false);
            p = q;
            continue;
        }
        if (!strncmp(p, "<!--", 4)) {
            const char* q = strstr(p + 4, "-->");
            if (!q) return fail("unterminated comment");
            p = q + 3;
        } else if (!strncmp(p, "<![CDATA[", 9)) {
            const char* q = strstr(p + 9, "]]>");
            if (!q) return fail("unterminated CDATA");
            add_text(p + 9, q, true);
            p = q + 3;
        } else if (p[1] == '?') {
            const char* q = strstr(p + 2, "?>");
            if (!q) return fail("unterminated processing instruction");
            p = q + 2;
        } else if (p[1] == '!') {
            const char* q = (const char*)memchr(p, '>', end - p);
            if (!q) return fail("unterminated declaration");
            p = q + 1;
        } else if (p[1] == '/') {
            const char* q = (const char*)memchr(p, '>', end - p);
            if (!q) return fail("unterminated end tag");
            std::string name(p + 2, q);
            while (!name.empty() && (name.back() == ' ' || name.back() == '\t' ||
                                     name.back() == '\n' || name.back() == '\r'))
                name.pop_back();
            if (stack.empty() || doc.nodes[stack.back()].name != name)
                return fail("mismatched end tag");
            stack.pop_back();
            p = q + 1;
        } else {
            const char* q = p + 1;
            const char* nb = q;
            while (q < end && !strchr(" \t\r\n/>", *q)) q++;
            if (q == nb) return fail("empty element name");
            XmlNode node;
            node.name.assign(nb, q);
            bool self_close = false;
            for (;;) {
                while (q < end && strchr(" \t\r\n", *q)) q++;
                if (q >= end) return fail("unterminated start tag");
                if (*q == '>') {
                    q++;
                    break;
                }
                if (*q == '/' && q + 1 < end && q[1] == '>') {
                    self_close = true;
                    q += 2;
                    break;
                }
                const char* an = q;
                while (q < end && !strchr(" \t\r\n=/>", *q)) q++;
                std::string aname(an, q);
                while (q < end && strchr(" \t\r\n", *q)) q++;
                if (q >= end || *q != '=') return fail("attribute without value");
                q++;
                while (q < end && strchr(" \t\r\n", *q)) q++;
                if (q >= end || (*q != '"' && *q != '\'')) return fail("unquoted attribute");
                const char quote = *q++;
                const char* vb = q;
                while (q < end && *q != quote) q++;
                if (q >= end) return fail("unterminated attribute");
                node.attrs.emplace_back(aname, decode_entities(vb, q));
                q++;
            }
            const int id = (int)doc.nodes.size();
            if (!stack.empty()) {
                XmlNode& par = doc.nodes[stack.back()];
                node.parent = stack.back();
                node.index_in_parent = (int)par.kids.size();
                par.kids.push_back(id);
            } else if (id != 0) {
                return fail("second root element");
            }
            doc.nodes.push_back(std::move(node));
            if (!self_close) stack.push_back(id);
            p = q;
        }
    }
    if (!stack.empty()) return fail("unclosed element");
    if (doc.nodes.empty()) return fail("no root element");
    return true;
}

int load_xml(const char* path, XmlDoc& doc, const char* who)
{
    if (!path) return set_error(TOPS_ERR_ARG, "%s: null path", who);
    FILE* f = fopen(path, "rb");
    if (!f) return set_error(TOPS_ERR_ARG, "%s: cannot open %s", who, path);
    std::string s;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, n);
    fclose(f);
    if (!parse_xml(s, doc))
        return set_error(TOPS_ERR_ARG, "%s: %s: %s", who, path, doc.error.c_str());
    return TOPS_OK;
}

// element at a path of child names below n (nullptr-terminated list)
int path_of(const XmlDoc& d, int n, std::initializer_list<const char*> names)
{
    for (const char* nm : names) n = d.child(n, nm);
    return n;
}

// getIntArray / getDoubleArray (Func.cpp:637-655, 679-699): split on one
// separator character, atoi/atof of each piece (an empty piece reads 0)
template <typename T, typename F>
int split_values(const char* s, char gap, T* out, int cap, F conv)
{
    if (!s) return 0;
    int count = 0;
    std::string tok;
    for (const char* p = s; *p; p++) {
        tok += *p;
        if (*p == gap) {
            if (count < cap) out[count] = conv(tok.c_str());
            count++;
            tok.clear();
        }
    }
    if (count < cap) out[count] = conv(tok.c_str());
    return count + 1;
}

struct Poly {
    double aztime, t0, c0, c1, c2;
};

// computeDC (Func.cpp:1828-1866): linear interpolation of the estimates
// bracketing az_time; before the first / after the last, extrapolation from
// the first / last pair.  (The reference's after-the-last branch indexes one
// past the list; the last pair is its evident intent.  One estimate is used
// as it is.)
Poly interpolate_dc(double az_time, const std::vector<Poly>& in)
{
#pragma clang fp contract(off)
    const int n = (int)in.size();
    if (n == 1) {
        Poly r = in[0];
        r.aztime = az_time;
        return r;
    }
    int i0 = 0, i1 = 1;
    if (az_time < in[0].aztime) {
        i0 = 0;
        i1 = 1;
    } else if (az_time > in[n - 1].aztime) {
        i0 = n - 2;
        i1 = n - 1;
    } else {
        for (int i = 0; i < n - 1; i++)
            if (az_time >= in[i].aztime && az_time < in[i + 1].aztime) {
                i0 = i;
                i1 = i + 1;
                break;
            }
    }
    Poly r;
    r.aztime = az_time;
    r.t0 = in[i0].t0;
    const double mu = (az_time - in[i0].aztime) / (in[i1].aztime - in[i0].aztime);
    r.c0 = (1 - mu) * in[i0].c0 + mu * in[i1].c0;
    r.c1 = (1 - mu) * in[i0].c1 + mu * in[i1].c1;
    r.c2 = (1 - mu) * in[i0].c2 + mu * in[i1].c2;
    return r;
}

template <typename T>
T* alloc_array(size_t n)
{
    return (T*)calloc(n ? n : 1, sizeof(T));
}

}  // namespace
}  // namespace tops

using namespace tops;

extern "C" {

double tops_utc2seconds(const char* utc)
{
    int year = 0, month = 0, day = 0;
    double hour = 0, min = 0, sec = 0;
    if (!utc) return 0.0;
    sscanf(utc, "%d-%d-%dT%lf:%lf:%lf", &year, &month, &day, &hour, &min, &sec);
    return hour * 60 * 60 + min * 60 + sec;
}

void tops_annotation_free(tops_annotation* a)
{
    if (!a) return;
    void* ptrs[] = {a->burst_first_line_time, a->burst_last_line_time,
                    a->burst_first_valid_line_time, a->burst_last_valid_line_time,
                    a->first_valid_line, a->last_valid_line, a->first_valid_sample,
                    a->last_valid_sample, a->geo_az_time, a->geo_slant_range_time,
                    a->geo_latitude, a->geo_longitude, a->geo_incidence_angle, a->burst_ll,
                    a->dc, a->fm};
    for (void* p : ptrs) free(p);
    memset(a, 0, sizeof(*a));
}

int tops_annotation_read(const char* xml_path, tops_annotation* a)
{
#pragma clang fp contract(off)
    const char* who = "tops_annotation_read";
    if (!a) return set_error(TOPS_ERR_ARG, "%s: null output", who);
    memset(a, 0, sizeof(*a));
    XmlDoc d;
    int rc = load_xml(xml_path, d, who);
    if (rc) return rc;
    const int root = 0;
    const char* missing = nullptr;
    auto txt = [&](std::initializer_list<const char*> path, const char* name) -> const char* {
        const char* t = d.text(path_of(d, root, path));
        if (!t && !missing) missing = name;
        return t ? t : "0";
    };
    const char* pol = txt({"adsHeader", "polarisation"}, "adsHeader/polarisation");
    snprintf(a->polarisation, sizeof(a->polarisation), "%s", pol);
    a->num_lines = atoi(txt({"imageAnnotation", "imageInformation", "numberOfLines"},
                            "numberOfLines"));
    a->num_samples = atoi(txt({"imageAnnotation", "imageInformation", "numberOfSamples"},
                              "numberOfSamples"));
    const char* flu = txt({"imageAnnotation", "imageInformation", "productFirstLineUtcTime"},
                          "productFirstLineUtcTime");
    const char* llu = txt({"imageAnnotation", "imageInformation", "productLastLineUtcTime"},
                          "productLastLineUtcTime");
    snprintf(a->first_line_utc, sizeof(a->first_line_utc), "%s", flu);
    snprintf(a->last_line_utc, sizeof(a->last_line_utc), "%s", llu);
    a->first_line_time = tops_utc2seconds(flu);
    a->last_line_time = tops_utc2seconds(llu);
    a->az_time_interval = atof(txt({"imageAnnotation", "imageInformation", "azimuthTimeInterval"},
                                   "azimuthTimeInterval"));
    a->range_pixel_spacing = atof(txt({"imageAnnotation", "imageInformation", "rangePixelSpacing"},
                                      "rangePixelSpacing"));
    a->azimuth_pixel_spacing = atof(
        txt({"imageAnnotation", "imageInformation", "azimuthPixelSpacing"}, "azimuthPixelSpacing"));
    a->radar_frequency = atof(txt({"generalAnnotation", "productInformation", "radarFrequency"},
                                  "radarFrequency"));
    a->wavelength = (double)kSol / a->radar_frequency;
    a->range_sampling_rate = atof(
        txt({"generalAnnotation", "productInformation", "rangeSamplingRate"}, "rangeSamplingRate"));
    a->azimuth_steering_rate = atof(txt(
        {"generalAnnotation", "productInformation", "azimuthSteeringRate"}, "azimuthSteeringRate"));
    a->heading = atof(txt({"generalAnnotation", "productInformation", "platformHeading"},
                          "platformHeading"));
    a->ascending_node_time = tops_utc2seconds(
        txt({"imageAnnotation", "imageInformation", "ascendingNodeTime"}, "ascendingNodeTime"));
    a->prf = double(1.0) / a->az_time_interval;
    a->azimuth_bandwidth = atof(txt({"imageAnnotation", "processingInformation",
                                     "swathProcParamsList", "swathProcParams", "azimuthProcessing",
                                     "processingBandwidth"},
                                    "azimuthProcessing/processingBandwidth"));
    a->range_bandwidth = atof(txt({"imageAnnotation", "processingInformation",
                                   "swathProcParamsList", "swathProcParams", "rangeProcessing",
                                   "processingBandwidth"},
                                  "rangeProcessing/processingBandwidth"));
    a->slr_time_first_pixel =
        atof(txt({"imageAnnotation", "imageInformation", "slantRangeTime"}, "slantRangeTime")) /
        2.0;
    a->slr_time_last_pixel =
        a->slr_time_first_pixel + (a->num_samples - 1) * a->range_pixel_spacing / kSol;
    const int burst_list = path_of(d, root, {"swathTiming", "burstList"});
    a->n_bursts = d.first_attr(burst_list) ? atoi(d.first_attr(burst_list)) : 0;
    a->lines_per_burst = atoi(txt({"swathTiming", "linesPerBurst"}, "linesPerBurst"));
    a->samples_per_burst = atoi(txt({"swathTiming", "samplesPerBurst"}, "samplesPerBurst"));
    if (missing) return set_error(TOPS_ERR_ARG, "%s: %s: no %s", who, xml_path, missing);
    const int nb = a->n_bursts, lpb = a->lines_per_burst;
    if (nb <= 0 || lpb <= 0 || a->samples_per_burst <= 0 || !(a->az_time_interval > 0))
        return set_error(TOPS_ERR_ARG, "%s: %s: bad burst list (%d bursts of %d lines)", who,
                         xml_path, nb, lpb);

    // ---- bursts: times and valid lines (Func.cpp:330-406)
    a->burst_first_line_time = alloc_array<double>(nb);
    a->burst_last_line_time = alloc_array<double>(nb);
    a->burst_first_valid_line_time = alloc_array<double>(nb);
    a->burst_last_valid_line_time = alloc_array<double>(nb);
    a->first_valid_line = alloc_array<int>(nb);
    a->last_valid_line = alloc_array<int>(nb);
    a->first_valid_sample = alloc_array<int>((size_t)nb * lpb);
    a->last_valid_sample = alloc_array<int>((size_t)nb * lpb);
    int fvp = a->num_samples, lvp = 0;
    int e = d.child(burst_list, nullptr);
    for (int k = 0; k < nb; k++, e = d.next_sibling(e)) {
        const char* az = d.text(d.child(e, "azimuthTime"));
        const char* fvs = d.text(d.child(e, "firstValidSample"));
        const char* lvs = d.text(d.child(e, "lastValidSample"));
        if (e < 0 || !az || !fvs || !lvs) {
            tops_annotation_free(a);
            return set_error(TOPS_ERR_ARG, "%s: %s: burst %d incomplete", who, xml_path, k);
        }
        a->burst_first_line_time[k] = tops_utc2seconds(az);
        a->burst_last_line_time[k] =
            a->burst_first_line_time[k] + (lpb - 1) * a->az_time_interval;
        int* f = a->first_valid_sample + (size_t)k * lpb;
        int* l = a->last_valid_sample + (size_t)k * lpb;
        // entries the list does not give read as invalid
        std::fill(f, f + lpb, -1);
        std::fill(l, l + lpb, -1);
        split_values(fvs, ' ', f, lpb, [](const char* s) { return atoi(s); });
        split_values(lvs, ' ', l, lpb, [](const char* s) { return atoi(s); });
        // first valid line, and the count of valid lines after it (the
        // reference counts them rather than taking the last index)
        int first = -1, last = -1;
        for (int y = 0; y < lpb; y++) {
            if (f[y] == -1) continue;
            if (f[y] < fvp) fvp = f[y];
            if (first == -1) {
                first = y;
                last = y;
            } else {
                last++;
            }
        }
        for (int y = 0; y < lpb; y++)
            if (l[y] != -1 && l[y] > lvp) lvp = l[y];
        a->burst_first_valid_line_time[k] =
            a->burst_first_line_time[k] + first * a->az_time_interval;
        a->burst_last_valid_line_time[k] =
            a->burst_first_line_time[k] + last * a->az_time_interval;
        a->first_valid_line[k] = first;
        a->last_valid_line[k] = last;
    }
    a->first_valid_pixel = fvp;
    a->last_valid_pixel = lvp;
    a->slr_time_first_valid_pixel =
        a->slr_time_first_pixel + a->first_valid_pixel * a->range_pixel_spacing / kSol;
    a->slr_time_last_valid_pixel =
        a->slr_time_first_pixel + a->last_valid_pixel * a->range_pixel_spacing / kSol;

    // ---- geolocation grid (Func.cpp:433-556)
    const int grid = path_of(d, root, {"geolocationGrid", "geolocationGridPointList"});
    const int npts = d.first_attr(grid) ? atoi(d.first_attr(grid)) : 0;
    if (npts <= 0) {
        tops_annotation_free(a);
        return set_error(TOPS_ERR_ARG, "%s: %s: no geolocation grid points", who, xml_path);
    }
    int per_line = 0;
    {
        // points of the first grid line (its line number is 0 in the
        // annotations; the reference counts the leading points of line 0)
        int g = d.child(grid, nullptr);
        const char* l0 = d.text(d.child(g, "line"));
        const int line0 = l0 ? atoi(l0) : 0;
        for (int k = 0; k < npts && g >= 0; k++, g = d.next_sibling(g)) {
            const char* ls = d.text(d.child(g, "line"));
            if (!ls || atoi(ls) != line0) break;
            per_line++;
        }
    }
    if (per_line <= 0) {
        tops_annotation_free(a);
        return set_error(TOPS_ERR_ARG, "%s: %s: bad geolocation grid", who, xml_path);
    }
    a->geo_points_per_line = per_line;
    a->geo_lines = npts / per_line;
    a->geo_az_time = alloc_array<double>(npts);
    a->geo_slant_range_time = alloc_array<double>(npts);
    a->geo_latitude = alloc_array<double>(npts);
    a->geo_longitude = alloc_array<double>(npts);
    a->geo_incidence_angle = alloc_array<double>(npts);
    {
        int g = d.child(grid, nullptr);
        for (int k = 0; k < npts; k++, g = d.next_sibling(g)) {
            const char* at = d.text(d.child(g, "azimuthTime"));
            const char* st = d.text(d.child(g, "slantRangeTime"));
            const char* la = d.text(d.child(g, "latitude"));
            const char* lo = d.text(d.child(g, "longitude"));
            const char* ia = d.text(d.child(g, "incidenceAngle"));
            if (g < 0 || !at || !st || !la || !lo || !ia) {
                tops_annotation_free(a);
                return set_error(TOPS_ERR_ARG, "%s: %s: geolocation point %d incomplete", who,
                                 xml_path, k);
            }
            a->geo_az_time[k] = tops_utc2seconds(at);
            a->geo_slant_range_time[k] = atof(st) / 2.0;
            a->geo_latitude[k] = atof(la);
            a->geo_longitude[k] = atof(lo);
            a->geo_incidence_angle[k] = atof(ia);
        }
    }
    a->near_range_on_left = npts > 1 && a->geo_incidence_angle[0] < a->geo_incidence_angle[1];
    if (a->geo_lines < nb + 1) {
        tops_annotation_free(a);
        return set_error(TOPS_ERR_ARG, "%s: %s: %d grid lines for %d bursts", who, xml_path,
                         a->geo_lines, nb);
    }
    a->burst_ll = alloc_array<double>((size_t)nb * 4);
    for (int b = 0; b < nb; b++) {
        const int P = per_line;
        const double la[4] = {a->geo_latitude[b * P], a->geo_latitude[b * P + P - 1],
                              a->geo_latitude[(b + 1) * P], a->geo_latitude[(b + 1) * P + P - 1]};
        const double lo[4] = {a->geo_longitude[b * P], a->geo_longitude[b * P + P - 1],
                              a->geo_longitude[(b + 1) * P],
                              a->geo_longitude[(b + 1) * P + P - 1]};
        a->burst_ll[4 * b + 0] = *std::min_element(la, la + 4);
        a->burst_ll[4 * b + 1] = *std::max_element(la, la + 4);
        a->burst_ll[4 * b + 2] = *std::min_element(lo, lo + 4);
        a->burst_ll[4 * b + 3] = *std::max_element(lo, lo + 4);
    }

    // ---- Doppler centroid estimates (Func.cpp:1887-1956)
    const char* method = d.text(path_of(d, root, {"imageAnnotation", "processingInformation",
                                                 "dcMethod"}));
    const int dcl = path_of(d, root, {"dopplerCentroid", "dcEstimateList"});
    const int ndc = d.first_attr(dcl) ? atoi(d.first_attr(dcl)) : 0;
    if (!method || ndc <= 0) {
        tops_annotation_free(a);
        return set_error(TOPS_ERR_ARG, "%s: %s: no Doppler centroid estimates", who, xml_path);
    }
    const bool data_analysis = !strcmp(method, "Data Analysis");
    std::vector<Poly> dc(ndc);
    {
        int g = d.child(dcl, nullptr);
        for (int k = 0; k < ndc; k++, g = d.next_sibling(g)) {
            const char* at = d.text(d.child(g, "azimuthTime"));
            const char* t0 = d.text(d.child(g, "t0"));
            const char* poly =
                d.text(d.child(g, data_analysis ? "dataDcPolynomial" : "geometryDcPolynomial"));
            if (g < 0 || !at || !t0 || !poly) {
                tops_annotation_free(a);
                return set_error(TOPS_ERR_ARG, "%s: %s: DC estimate %d incomplete", who, xml_path,
                                 k);
            }
            double c[3] = {0, 0, 0};
            split_values(poly, ' ', c, 3, [](const char* s) { return atof(s); });
            dc[k] = {tops_utc2seconds(at), atof(t0), c[0], c[1], c[2]};
        }
    }
    a->dc = alloc_array<double>((size_t)nb * 4);
    for (int b = 0; b < nb; b++) {
        Poly p;
        if (ndc > nb) {
            p = dc[b];
        } else {
            const double centre =
                0.5 * (a->burst_first_line_time[b] + a->burst_last_line_time[b]);
            p = interpolate_dc(centre, dc);
        }
        a->dc[4 * b + 0] = p.t0;
        a->dc[4 * b + 1] = p.c0;
        a->dc[4 * b + 2] = p.c1;
        a->dc[4 * b + 3] = p.c2;
    }

    // ---- azimuth FM rates (Func.cpp:1972-2014): burst b takes estimate b
    const int fml = path_of(d, root, {"generalAnnotation", "azimuthFmRateList"});
    const int nfm = d.first_attr(fml) ? atoi(d.first_attr(fml)) : 0;
    if (nfm < nb) {
        tops_annotation_free(a);
        return set_error(TOPS_ERR_ARG, "%s: %s: %d FM-rate estimates for %d bursts", who,
                         xml_path, nfm, nb);
    }
    a->fm = alloc_array<double>((size_t)nb * 4);
    {
        int g = d.child(fml, nullptr);
        for (int k = 0; k < nb; k++, g = d.next_sibling(g)) {
            const char* t0 = d.text(d.child(g, "t0"));
            double c[3] = {0, 0, 0};
            const char* poly = d.text(d.child(g, "azimuthFmRatePolynomial"));
            bool ok = g >= 0 && t0;
            if (ok && d.child(g, "azimuthFmRatePolynomial") >= 0) {
                ok = poly != nullptr;
                if (ok) split_values(poly, ' ', c, 3, [](const char* s) { return atof(s); });
            } else if (ok) {
                const char* c0 = d.text(d.child(g, "c0"));
                const char* c1 = d.text(d.child(g, "c1"));
                const char* c2 = d.text(d.child(g, "c2"));
                ok = c0 && c1 && c2;
                if (ok) {
                    c[0] = atof(c0);
                    c[1] = atof(c1);
                    c[2] = atof(c2);
                }
            }
            if (!ok) {
                tops_annotation_free(a);
                return set_error(TOPS_ERR_ARG, "%s: %s: FM-rate estimate %d incomplete", who,
                                 xml_path, k);
            }
            a->fm[4 * k + 0] = atof(t0);
            a->fm[4 * k + 1] = c[0];
            a->fm[4 * k + 2] = c[1];
            a->fm[4 * k + 3] = c[2];
        }
    }
    return TOPS_OK;
}

// ---------------------------------------------------------------------------
// EOF precise orbit (Func.cpp:704-819)
// ---------------------------------------------------------------------------
void tops_eof_free(tops_eof_orbit* o)
{
    if (!o) return;
    void* ptrs[] = {o->t, o->x, o->y, o->z, o->vx, o->vy, o->vz};
    for (void* p : ptrs) free(p);
    memset(o, 0, sizeof(*o));
}

int tops_eof_read(const char* eof_path, const char* first_utc, const char* last_utc,
                  tops_eof_orbit* o)
{
    const char* who = "tops_eof_read";
    if (!o || !first_utc || !last_utc) return set_error(TOPS_ERR_ARG, "%s: null argument", who);
    memset(o, 0, sizeof(*o));
    int y0 = 0, mo0 = 0, d0 = 0, h0 = 0, mi0 = 0, s0 = 0;
    int y1 = 0, mo1 = 0, d1 = 0, h1 = 0, mi1 = 0, s1 = 0;
    if (sscanf(first_utc, "%d-%d-%dT%d:%d:%d", &y0, &mo0, &d0, &h0, &mi0, &s0) != 6 ||
        sscanf(last_utc, "%d-%d-%dT%d:%d:%d", &y1, &mo1, &d1, &h1, &mi1, &s1) != 6)
        return set_error(TOPS_ERR_ARG, "%s: bad UTC time", who);
    // the 10 s slot of the first time ("...T12:34:5" matches 12:34:50-59)
    char slot[64];
    snprintf(slot, sizeof(slot), "%04d-%02d-%02dT%02d:%02d:%01d", y0, mo0, d0, h0, mi0, s0 / 10);
    // DiffUTCTime (Func.cpp:657-677): whole seconds, one day added when the
    // dates differ
    int tt = h1 * 3600 + mi1 * 60 + s1 - (h0 * 3600 + mi0 * 60 + s0);
    if (!(y0 == y1 && mo0 == mo1 && d0 == d1)) tt += 24 * 3600;
    const int n = tt / 10 + 10;
    XmlDoc d;
    int rc = load_xml(eof_path, d, who);
    if (rc) return rc;
    const int list = path_of(d, 0, {"Data_Block", "List_of_OSVs"});
    std::vector<int> osv;
    for (int g = d.child(list, nullptr); g >= 0; g = d.next_sibling(g)) osv.push_back(g);
    int index = -1;
    for (size_t k = 0; k < osv.size(); k++) {
        const char* u = d.text(d.child(osv[k], "UTC"));
        if (u && strstr(u, slot)) {
            index = (int)k;
            break;
        }
    }
    if (index < 0)
        return set_error(TOPS_ERR_ARG, "%s: %s: no state vector in the slot %s", who, eof_path,
                         slot);
    const int start = index - 4;
    if (n <= 1 || start < 0 || start + n > (int)osv.size())
        return set_error(TOPS_ERR_ARG, "%s: %s: %d state vectors from %d needed, %d in the file",
                         who, eof_path, n, start, (int)osv.size());
    o->n = n;
    double** arrs[] = {&o->t, &o->x, &o->y, &o->z, &o->vx, &o->vy, &o->vz};
    for (double** p : arrs) *p = alloc_array<double>(n);
    static const char* tags[] = {"X", "Y", "Z", "VX", "VY", "VZ"};
    for (int i = 0; i < n; i++) {
        const int g = osv[start + i];
        const char* u = d.text(d.child(g, "UTC"));
        if (!u || strlen(u) < 5) {
            tops_eof_free(o);
            return set_error(TOPS_ERR_ARG, "%s: state vector %d has no UTC", who, start + i);
        }
        // "UTC=..." without its prefix (the reference's substr(4, 30))
        std::string t(u);
        t = t.substr(4, 30);
        o->t[i] = tops_utc2seconds(t.c_str());
        for (int c = 0; c < 6; c++) {
            const char* v = d.text(d.child(g, tags[c]));
            if (!v) {
                tops_eof_free(o);
                return set_error(TOPS_ERR_ARG, "%s: state vector %d has no %s", who, start + i,
                                 tags[c]);
            }
            (*arrs[c + 1])[i] = atof(v);
        }
    }
    o->time_interval = (o->t[n - 1] - o->t[0]) / (n - 1);
    return TOPS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// TIFF
// ---------------------------------------------------------------------------
struct tops_tiff {
    int fd = -1;
    bool writable = false;
    bool big_endian = false;
    int width = 0, height = 0, data_type = 0, elem = 0, comp = 0;   // comp: component bytes
    int rows_per_strip = 0;
    std::vector<uint64_t> strip_offsets;
    // device staging (tops_tiff_read_device)
    void* pinned[2] = {nullptr, nullptr};
    size_t pinned_bytes = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
};

namespace tops {
namespace {

int type_layout(int data_type, int* elem, int* comp, int* bits, int* fmt)
{
    switch (data_type) {
    case TOPS_GDT_CINT16: *elem = 4; *comp = 2; *bits = 32; *fmt = 5; return 1;
    case TOPS_GDT_FLOAT32: *elem = 4; *comp = 4; *bits = 32; *fmt = 3; return 1;
    case TOPS_GDT_FLOAT64: *elem = 8; *comp = 8; *bits = 64; *fmt = 3; return 1;
    case TOPS_GDT_CFLOAT32: *elem = 8; *comp = 4; *bits = 64; *fmt = 6; return 1;
    default: return 0;
    }
}

int type_of(int fmt, int bits)
{
    if (fmt == 5 && bits == 32) return TOPS_GDT_CINT16;
    if (fmt == 3 && bits == 32) return TOPS_GDT_FLOAT32;
    if (fmt == 3 && bits == 64) return TOPS_GDT_FLOAT64;
    if (fmt == 6 && bits == 64) return TOPS_GDT_CFLOAT32;
    return 0;
}

bool read_at(int fd, void* buf, size_t n, uint64_t off)
{
    char* p = (char*)buf;
    while (n > 0) {
        const ssize_t r = pread(fd, p, n, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        off += (uint64_t)r;
    }
    return true;
}

bool write_at(int fd, const void* buf, size_t n, uint64_t off)
{
    const char* p = (const char*)buf;
    while (n > 0) {
        const ssize_t r = pwrite(fd, p, n, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        off += (uint64_t)r;
    }
    return true;
}

uint64_t get_u(const unsigned char* p, int n, bool be)
{
    uint64_t v = 0;
    for (int i = 0; i < n; i++) v |= (uint64_t)p[be ? n - 1 - i : i] << (8 * i);
    return v;
}

void swap_components(void* buf, size_t bytes, int comp)
{
    if (comp <= 1) return;
    unsigned char* p = (unsigned char*)buf;
    for (size_t i = 0; i + comp <= bytes; i += comp) std::reverse(p + i, p + i + comp);
}

// the values of one IFD entry as unsigned integers
bool entry_values(int fd, bool be, bool big, const unsigned char* ent, std::vector<uint64_t>& out)
{
    const int type = (int)get_u(ent + 2, 2, be);
    const uint64_t count = get_u(ent + 4, big ? 8 : 4, be);
    int sz;
    switch (type) {
    case 1: sz = 1; break;    // BYTE
    case 3: sz = 2; break;    // SHORT
    case 4: sz = 4; break;    // LONG
    case 16: sz = 8; break;   // LONG8
    default: return false;
    }
    if (count > (1u << 28)) return false;
    const int field = big ? 8 : 4;
    const unsigned char* vp = ent + (big ? 12 : 8);
    std::vector<unsigned char> ext;
    if (count * sz > (uint64_t)field) {
        ext.resize(count * sz);
        if (!read_at(fd, ext.data(), ext.size(), get_u(vp, field, be))) return false;
        vp = ext.data();
    }
    out.resize(count);
    for (uint64_t i = 0; i < count; i++) out[i] = get_u(vp + i * sz, sz, be);
    return true;
}

int check_window(const tops_tiff* t, int x0, int y0, int lines, int pixels, int data_type,
                 const char* who)
{
    if (!t || t->fd < 0) return set_error(TOPS_ERR_ARG, "%s: no open TIFF", who);
    if (data_type != t->data_type)
        return set_error(TOPS_ERR_ARG, "%s: raster type %d, requested %d", who, t->data_type,
                         data_type);
    if (x0 < 0 || y0 < 0 || lines < 0 || pixels < 0 || x0 + pixels > t->width ||
        y0 + lines > t->height)
        return set_error(TOPS_ERR_ARG, "%s: window (%d, %d) %d x %d outside %d x %d", who, x0, y0,
                         pixels, lines, t->width, t->height);
    return TOPS_OK;
}

uint64_t row_offset(const tops_tiff* t, int row, int x0)
{
    const int s = row / t->rows_per_strip;
    return t->strip_offsets[s] +
           ((uint64_t)(row % t->rows_per_strip) * t->width + x0) * (uint64_t)t->elem;
}

// rows y0 .. y0+lines-1, columns x0 .. x0+pixels-1 into buf; runs of rows
// that are adjacent in the file are read with one call
int read_rows(tops_tiff* t, int x0, int y0, int lines, int pixels, void* buf)
{
    const size_t rb = (size_t)pixels * t->elem;
    int r = 0;
    while (r < lines) {
        const uint64_t off = row_offset(t, y0 + r, x0);
        int n = 1;
        while (r + n < lines && row_offset(t, y0 + r + n, x0) == off + (uint64_t)n * rb) n++;
        char* dst = (char*)buf + (size_t)r * rb;
        if (!read_at(t->fd, dst, rb * n, off))
            return set_error(TOPS_ERR_ARG, "tops_tiff_read: short read at row %d", y0 + r);
        if (t->big_endian) swap_components(dst, rb * n, t->comp);
        r += n;
    }
    return TOPS_OK;
}

}  // namespace
}  // namespace tops

extern "C" {

tops_tiff* tops_tiff_open(const char* path)
{
    const char* who = "tops_tiff_open";
    if (!path) {
        set_error(TOPS_ERR_ARG, "%s: null path", who);
        return nullptr;
    }
    const int fd = open(path, O_RDONLY);
    if (fd < 0) {
        set_error(TOPS_ERR_ARG, "%s: cannot open %s", who, path);
        return nullptr;
    }
    auto fail = [&](const char* what) -> tops_tiff* {
        close(fd);
        set_error(TOPS_ERR_ARG, "%s: %s: %s", who, path, what);
        return nullptr;
    };
    unsigned char h[16];
    if (!read_at(fd, h, 8, 0)) return fail("short header");
    bool be;
    if (h[0] == 'I' && h[1] == 'I') be = false;
    else if (h[0] == 'M' && h[1] == 'M') be = true;
    else return fail("not a TIFF");
    const int magic = (int)get_u(h + 2, 2, be);
    bool big;
    uint64_t ifd;
    if (magic == 42) {
        big = false;
        ifd = get_u(h + 4, 4, be);
    } else if (magic == 43) {
        big = true;
        if (!read_at(fd, h, 16, 0)) return fail("short BigTIFF header");
        if (get_u(h + 4, 2, be) != 8) return fail("BigTIFF offset size is not 8");
        ifd = get_u(h + 8, 8, be);
    } else {
        return fail("not a TIFF");
    }
    unsigned char nb[8];
    if (!read_at(fd, nb, big ? 8 : 2, ifd)) return fail("short IFD");
    const uint64_t nent = get_u(nb, big ? 8 : 2, be);
    const int esz = big ? 20 : 12;
    if (nent == 0 || nent > 4096) return fail("bad IFD");
    std::vector<unsigned char> ents(nent * esz);
    if (!read_at(fd, ents.data(), ents.size(), ifd + (big ? 8 : 2))) return fail("short IFD");
    uint64_t width = 0, height = 0, bits = 0, compression = 1, spp = 1, rps = 0, planar = 1,
             fmt = 1;
    std::vector<uint64_t> offsets;
    bool tiled = false;
    for (uint64_t i = 0; i < nent; i++) {
        const unsigned char* e = &ents[i * esz];
        const int tag = (int)get_u(e, 2, be);
        std::vector<uint64_t> v;
        switch (tag) {
        case 256: case 257: case 258: case 259: case 273: case 277: case 278: case 284:
        case 339:
            if (!entry_values(fd, be, big, e, v) || v.empty()) return fail("bad tag value");
            break;
        case 322: case 323: case 324: case 325:
            tiled = true;
            break;
        default:
            break;
        }
        switch (tag) {
        case 256: width = v[0]; break;
        case 257: height = v[0]; break;
        case 258: bits = v[0]; break;
        case 259: compression = v[0]; break;
        case 273: offsets = v; break;
        case 277: spp = v[0]; break;
        case 278: rps = v[0]; break;
        case 284: planar = v[0]; break;
        case 339: fmt = v[0]; break;
        default: break;
        }
    }
    if (tiled) return fail("tiled TIFF (only stripped rasters are read)");
    if (compression != 1) return fail("compressed TIFF (only uncompressed rasters are read)");
    if (spp != 1 || planar != 1) return fail("more than one band");
    const int type = type_of((int)fmt, (int)bits);
    if (!type) return fail("unsupported sample type");
    if (width == 0 || height == 0 || width > (1u << 30) || height > (1u << 30))
        return fail("bad raster size");
    if (rps == 0 || rps > height) rps = height;
    const uint64_t nstrips = (height + rps - 1) / rps;
    if (offsets.size() < nstrips) return fail("missing strip offsets");
    tops_tiff* t = new tops_tiff();
    t->fd = fd;
    t->big_endian = be;
    t->width = (int)width;
    t->height = (int)height;
    t->data_type = type;
    int b2, f2;
    type_layout(type, &t->elem, &t->comp, &b2, &f2);
    t->rows_per_strip = (int)rps;
    t->strip_offsets.assign(offsets.begin(), offsets.begin() + nstrips);
    return t;
}

int tops_tiff_info(const tops_tiff* t, int* pixels, int* lines, int* data_type)
{
    if (!t) return set_error(TOPS_ERR_ARG, "tops_tiff_info: null handle");
    if (pixels) *pixels = t->width;
    if (lines) *lines = t->height;
    if (data_type) *data_type = t->data_type;
    return TOPS_OK;
}

int tops_tiff_read(tops_tiff* t, int x0, int y0, int lines, int pixels, int data_type,
                   void* buffer)
{
    int rc = check_window(t, x0, y0, lines, pixels, data_type, "tops_tiff_read");
    if (rc) return rc;
    if (!buffer) return set_error(TOPS_ERR_ARG, "tops_tiff_read: null buffer");
    return read_rows(t, x0, y0, lines, pixels, buffer);
}

int tops_tiff_read_overlaps(tops_tiff* t, void* data_for, void* data_back, int n_overlap,
                            const int* overlap_size, int lines_per_burst, int x0, int y0,
                            int width, int data_type)
{
    const char* who = "tops_tiff_read_overlaps";
    if (!t || !data_for || !data_back || (n_overlap > 0 && !overlap_size))
        return set_error(TOPS_ERR_ARG, "%s: null argument", who);
    size_t off = 0;
    for (int i = 0; i < n_overlap; i++) {
        const int n = overlap_size[i];
        const int yf = y0 + (i + 1) * lines_per_burst - n, yb = y0 + (i + 1) * lines_per_burst;
        int rc = check_window(t, x0, yf, n, width, data_type, who);
        if (!rc) rc = check_window(t, x0, yb, n, width, data_type, who);
        if (rc) return rc;
        rc = read_rows(t, x0, yf, n, width, (char*)data_for + off * t->elem);
        if (!rc) rc = read_rows(t, x0, yb, n, width, (char*)data_back + off * t->elem);
        if (rc) return rc;
        off += (size_t)n * width;
    }
    return TOPS_OK;
}

int tops_tiff_read_device(tops_tiff* t, int x0, int y0, int lines, int pixels, int data_type,
                          void* d_buffer, void* stream)
{
    const char* who = "tops_tiff_read_device";
    int rc = check_window(t, x0, y0, lines, pixels, data_type, who);
    if (rc) return rc;
    if (!d_buffer) return set_error(TOPS_ERR_ARG, "%s: null buffer", who);
    const size_t rb = (size_t)pixels * t->elem;
    if (rb == 0 || lines == 0) return TOPS_OK;
    constexpr size_t kStage = 32u << 20;   // bytes per staging buffer
    if (!t->pinned[0]) {
        for (int i = 0; i < 2; i++) {
            TOPS_HIP(hipHostMalloc(&t->pinned[i], kStage, hipHostMallocDefault));
            TOPS_HIP(hipEventCreateWithFlags(&t->ev[i], hipEventDisableTiming));
        }
        t->pinned_bytes = kStage;
    }
    const int chunk = (int)std::max<size_t>(1, t->pinned_bytes / rb);
    if ((size_t)chunk * rb > t->pinned_bytes)
        return set_error(TOPS_ERR_ARG, "%s: a row of %zu bytes exceeds the staging buffer", who,
                         rb);
    hipStream_t s = (hipStream_t)stream;
    bool used[2] = {false, false};
    int k = 0;
    for (int r = 0; r < lines; r += chunk, k ^= 1) {
        const int n = std::min(chunk, lines - r);
        // the copy that last read this staging buffer must be done
        if (used[k]) TOPS_HIP(hipEventSynchronize(t->ev[k]));
        rc = read_rows(t, x0, y0 + r, n, pixels, t->pinned[k]);
        if (rc) break;
        TOPS_HIP(hipMemcpyAsync((char*)d_buffer + (size_t)r * rb, t->pinned[k], rb * n,
                                hipMemcpyHostToDevice, s));
        TOPS_HIP(hipEventRecord(t->ev[k], s));
        used[k] = true;
    }
    for (int i = 0; i < 2; i++)
        if (used[i]) TOPS_HIP(hipEventSynchronize(t->ev[i]));
    return rc;
}

tops_tiff* tops_tiff_create(const char* path, int data_type, int pixels, int lines)
{
    const char* who = "tops_tiff_create";
    int elem, comp, bits, fmt;
    if (!path || !type_layout(data_type, &elem, &comp, &bits, &fmt)) {
        set_error(TOPS_ERR_ARG, "%s: unsupported data type %d", who, data_type);
        return nullptr;
    }
    if (pixels <= 0 || lines <= 0) {
        set_error(TOPS_ERR_ARG, "%s: create image error, width = %d height = %d", who, pixels,
                  lines);
        return nullptr;
    }
    const uint64_t data_bytes = (uint64_t)pixels * lines * elem;
    const bool big = data_bytes + 16 * (uint64_t)lines + 4096 >= 0xFFFFFFFFull;
    // header | IFD (11 entries) | strip offsets | strip byte counts | rows
    const int nent = 11, esz = big ? 20 : 12, osz = big ? 8 : 4;
    const uint64_t ifd = big ? 16 : 8;
    const uint64_t ifd_bytes = (big ? 8 : 2) + (uint64_t)nent * esz + (big ? 8 : 4);
    const uint64_t off_arr = ifd + ifd_bytes;
    const uint64_t cnt_arr = off_arr + (uint64_t)lines * osz;
    const uint64_t data0 = (cnt_arr + (uint64_t)lines * osz + 15) / 16 * 16;
    const int fd = open(path, O_RDWR | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) {
        set_error(TOPS_ERR_ARG, "%s: cannot create %s", who, path);
        return nullptr;
    }
    std::vector<unsigned char> hdr(data0, 0);
    auto put = [&](uint64_t at, uint64_t v, int n) {
        for (int i = 0; i < n; i++) hdr[at + i] = (unsigned char)(v >> (8 * i));
    };
    hdr[0] = hdr[1] = 'I';
    if (big) {
        put(2, 43, 2);
        put(4, 8, 2);
        put(8, ifd, 8);
    } else {
        put(2, 42, 2);
        put(4, ifd, 4);
    }
    put(ifd, nent, big ? 8 : 2);
    uint64_t e = ifd + (big ? 8 : 2);
    const int long_type = big ? 16 : 4;
    auto entry = [&](int tag, int type, uint64_t count, uint64_t value) {
        put(e, tag, 2);
        put(e + 2, type, 2);
        put(e + 4, count, big ? 8 : 4);
        const int tsz = type == 3 ? 2 : type == 4 ? 4 : 8;
        put(e + (big ? 12 : 8), value, count == 1 ? tsz : osz);
        e += esz;
    };
    entry(256, 4, 1, (uint64_t)pixels);                   // ImageWidth
    entry(257, 4, 1, (uint64_t)lines);                    // ImageLength
    entry(258, 3, 1, (uint64_t)bits);                     // BitsPerSample
    entry(259, 3, 1, 1);                                  // Compression: none
    entry(262, 3, 1, 1);                                  // Photometric: BlackIsZero
    entry(273, long_type, (uint64_t)lines, lines == 1 ? data0 : off_arr);   // StripOffsets
    entry(277, 3, 1, 1);                                  // SamplesPerPixel
    entry(278, 4, 1, 1);                                  // RowsPerStrip
    entry(279, long_type, (uint64_t)lines,
          lines == 1 ? (uint64_t)pixels * elem : cnt_arr);  // StripByteCounts
    entry(284, 3, 1, 1);                                  // PlanarConfiguration
    entry(339, 3, 1, (uint64_t)fmt);                      // SampleFormat
    put(e, 0, big ? 8 : 4);                               // no next IFD
    tops_tiff* t = new tops_tiff();
    t->fd = fd;
    t->writable = true;
    t->width = pixels;
    t->height = lines;
    t->data_type = data_type;
    t->elem = elem;
    t->comp = comp;
    t->rows_per_strip = 1;
    t->strip_offsets.resize(lines);
    const uint64_t rb = (uint64_t)pixels * elem;
    for (int r = 0; r < lines; r++) t->strip_offsets[r] = data0 + (uint64_t)r * rb;
    bool ok = true;
    if (lines > 1) {
        for (int r = 0; r < lines; r++) {
            put(off_arr + (uint64_t)r * osz, t->strip_offsets[r], osz);
            put(cnt_arr + (uint64_t)r * osz, rb, osz);
        }
    }
    ok = write_at(fd, hdr.data(), hdr.size(), 0) &&
         ftruncate(fd, (off_t)(data0 + data_bytes)) == 0;
    if (!ok) {
        set_error(TOPS_ERR_ARG, "%s: cannot write %s", who, path);
        close(fd);
        delete t;
        return nullptr;
    }
    return t;
}

int tops_tiff_write(tops_tiff* t, int x0, int y0, int lines, int pixels, int data_type,
                    const void* buffer)
{
    const char* who = "tops_tiff_write";
    int rc = check_window(t, x0, y0, lines, pixels, data_type, who);
    if (rc) return rc;
    if (!t->writable) return set_error(TOPS_ERR_ARG, "%s: TIFF opened for reading", who);
    if (!buffer) return set_error(TOPS_ERR_ARG, "%s: null buffer", who);
    const size_t rb = (size_t)pixels * t->elem;
    if (pixels == t->width) {
        // whole rows are adjacent in the file
        if (!write_at(t->fd, buffer, rb * lines, row_offset(t, y0, 0)))
            return set_error(TOPS_ERR_ARG, "%s: write failed", who);
        return TOPS_OK;
    }
    for (int r = 0; r < lines; r++)
        if (!write_at(t->fd, (const char*)buffer + (size_t)r * rb, rb, row_offset(t, y0 + r, x0)))
            return set_error(TOPS_ERR_ARG, "%s: write failed", who);
    return TOPS_OK;
}

int tops_tiff_close(tops_tiff* t)
{
    if (!t) return TOPS_OK;
    int rc = TOPS_OK;
    if (t->fd >= 0 && close(t->fd) != 0) rc = set_error(TOPS_ERR_ARG, "tops_tiff_close: close failed");
    for (int i = 0; i < 2; i++) {
        if (t->ev[i]) (void)hipEventDestroy(t->ev[i]);
        if (t->pinned[i]) (void)hipHostFree(t->pinned[i]);
    }
    delete t;
    return rc;
}

}  // extern "C"
