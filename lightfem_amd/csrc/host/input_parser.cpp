// `.fem` reader with the reference's grammar and error texts
// (src/input_parser.cpp:162-342, docs/input_format.md): '#' comments, [section]
// headers with aliases, whitespace-separated fields, fatal unknown or repeated
// settings, every error reported as "<source>:<line>: <message>". Whole files take a
// parallel fast path for the node, material and element rows, and fall back to the
// line-by-line reader whenever a line is anything but plain decimal fields.
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <optional>
#include <set>
#include <sstream>

#include "lightfem/lightfem.hpp"

namespace lightfem {

namespace {

enum class Section { None, Settings, Nodes, Materials, Elements, Trusses, Boundary, Loads, BodyForces, EdgeLoads };

std::string lowercase(std::string s) {
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

std::string strip(const std::string& s) {
    std::size_t b = 0, e = s.size();
    while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
    while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
    return s.substr(b, e - b);
}

class LineReader {
public:
    LineReader(const std::string& source, std::size_t line) : source_(source), line_(line) {}

    [[noreturn]] void fail(const std::string& message) const { throw ParseError(source_, line_, message); }

    int integer(const std::string& token, const char* field) const {
        long long value = 0;
        std::size_t used = 0;
        try {
            value = std::stoll(token, &used);
        } catch (const std::invalid_argument&) {
            fail(std::string("invalid integer for ") + field + ": " + token);
        } catch (const std::out_of_range&) {
            fail(std::string("integer out of range for ") + field + ": " + token);
        }
        if (used != token.size() || value < std::numeric_limits<int>::min() || value > std::numeric_limits<int>::max())
            fail(std::string("invalid integer for ") + field + ": " + token);
        return static_cast<int>(value);
    }

    std::size_t count(const std::string& token, const char* field) const {
        unsigned long long value = 0;
        std::size_t used = 0;
        try {
            value = std::stoull(token, &used);
        } catch (const std::invalid_argument&) {
            fail(std::string("invalid non-negative integer for ") + field + ": " + token);
        } catch (const std::out_of_range&) {
            fail(std::string("integer out of range for ") + field + ": " + token);
        }
        if (used != token.size()) fail(std::string("invalid non-negative integer for ") + field + ": " + token);
        return static_cast<std::size_t>(value);
    }

    double real(const std::string& token, const char* field) const {
        double value = 0.0;
        std::size_t used = 0;
        try {
            value = std::stod(token, &used);
        } catch (const std::invalid_argument&) {
            fail(std::string("invalid number for ") + field + ": " + token);
        } catch (const std::out_of_range&) {
            fail(std::string("number out of range for ") + field + ": " + token);
        }
        if (used != token.size() || !std::isfinite(value))
            fail(std::string("invalid finite number for ") + field + ": " + token);
        return value;
    }

    std::optional<double> constraint(const std::string& token, const char* field) const {
        const std::string t = lowercase(token);
        if (t == "free" || t == "none" || t == "-") return std::nullopt;
        return real(token, field);
    }

    void expect(const std::vector<std::string>& tokens, std::size_t n, const char* format) const {
        if (tokens.size() != n)
            fail("expected " + std::to_string(n) + " fields (" + format + "), got " + std::to_string(tokens.size()));
    }

    std::size_t line() const { return line_; }

private:
    const std::string& source_;
    std::size_t line_;
};

Section section_named(const std::string& raw, const LineReader& at) {
    static const std::map<std::string, Section> names = {
        {"settings", Section::Settings},       {"nodes", Section::Nodes},
        {"materials", Section::Materials},     {"elements", Section::Elements},
        {"trusses", Section::Trusses},         {"bars", Section::Trusses},
        {"rods", Section::Trusses},            {"members", Section::Trusses},
        {"boundary", Section::Boundary},       {"boundaries", Section::Boundary},
        {"dirichlet", Section::Boundary},      {"loads", Section::Loads},
        {"forces", Section::Loads},            {"body_forces", Section::BodyForces},
        {"body_loads", Section::BodyForces},   {"edge_loads", Section::EdgeLoads},
        {"surface_loads", Section::EdgeLoads}, {"tractions", Section::EdgeLoads},
        {"pressures", Section::EdgeLoads}};
    const std::string name = lowercase(strip(raw));
    const auto it = names.find(name);
    if (it == names.end()) at.fail("unknown section [" + name + "]");
    return it->second;
}

void read_setting(Model& model, const std::vector<std::string>& tokens, const LineReader& at) {
    at.expect(tokens, 2, "key value");
    SolverSettings& s = model.settings;
    const std::string key = lowercase(tokens[0]);
    const std::string& value = tokens[1];
    using Setter = std::function<void()>;
    const std::map<std::string, Setter> setters = {
        {"plane_mode", [&] {
             const std::string mode = lowercase(value);
             if (mode == "plane_stress") s.plane_mode = PlaneMode::PlaneStress;
             else if (mode == "plane_strain") s.plane_mode = PlaneMode::PlaneStrain;
             else at.fail("plane_mode must be plane_stress or plane_strain");
         }},
        {"solver", [&] { s.solver = lowercase(value); }},
        {"tolerance", [&] { s.tolerance = at.real(value, "tolerance"); }},
        {"max_iterations", [&] { s.max_iterations = at.count(value, "max_iterations"); }},
        {"reordering", [&] { s.reordering = lowercase(value); }},
        {"assembly", [&] { s.assembly = lowercase(value); }},
        {"threads", [&] { s.threads = at.count(value, "threads"); }},
        {"storage", [&] { s.storage = lowercase(value); }},
        {"preconditioner", [&] { s.preconditioner = lowercase(value); }},
        {"subdomain_split", [&] { s.subdomain_split = lowercase(value); }},
        {"subdomain_nodes", [&] { s.subdomain_nodes = at.count(value, "subdomain_nodes"); }},
        {"output", [&] { s.output_file = value; }},
        {"deformation_scale", [&] { s.deformation_scale = at.real(value, "deformation_scale"); }},
        {"csv_prefix", [&] { s.csv_prefix = value; }},
        {"precision", [&] { s.precision = lowercase(value); }},
        {"stop_rule", [&] { s.stop_rule = lowercase(value); }},
    };
    const auto it = setters.find(key);
    if (it == setters.end()) at.fail("unknown setting: " + key);
    it->second();
}

// One data line of `section` (tokens t) into the model: input_parser.cpp:230-342.
void apply_line(Model& model, Section section, const std::vector<std::string>& t, const LineReader& at,
                std::set<std::string>& seen, std::size_t number) {
    switch (section) {
        case Section::Settings:
            if (t.size() == 2 && !seen.insert(lowercase(t[0])).second) at.fail("duplicate setting: " + lowercase(t[0]));
            read_setting(model, t, at);
            break;
        case Section::Nodes:
            at.expect(t, 3, "id x y");
            model.nodes.push_back(Node{at.integer(t[0], "node id"), Vec2{at.real(t[1], "x"), at.real(t[2], "y")},
                                       std::numeric_limits<std::size_t>::max(),
                                       std::numeric_limits<std::size_t>::max(), number});
            break;
        case Section::Materials:
            at.expect(t, 4, "id E nu thickness");
            model.materials.push_back(Material{at.integer(t[0], "material id"), at.real(t[1], "E"),
                                               at.real(t[2], "nu"), at.real(t[3], "thickness"), number});
            break;
        case Section::Elements:
            at.expect(t, 6, "id n1 n2 n3 n4 material_id");
            model.elements.push_back(Q4Element{at.integer(t[0], "element id"),
                                               {{at.integer(t[1], "n1"), at.integer(t[2], "n2"),
                                                 at.integer(t[3], "n3"), at.integer(t[4], "n4")}},
                                               at.integer(t[5], "material id"), number});
            break;
        case Section::Trusses:
            at.expect(t, 5, "id n1 n2 material_id area");
            model.trusses.push_back(TrussElement{at.integer(t[0], "truss id"),
                                                 {{at.integer(t[1], "n1"), at.integer(t[2], "n2")}},
                                                 at.integer(t[3], "material id"), at.real(t[4], "area"), number});
            break;
        case Section::Boundary:
            if (t.size() == 2) {
                const std::string kind = lowercase(t[1]);
                const int node = at.integer(t[0], "boundary node id");
                if (kind == "fixed" || kind == "pinned") {
                    model.boundaries.push_back(DirichletBC{node, 0.0, 0.0, number});
                } else if (kind == "roller_x" || kind == "symmetry_x") {
                    model.boundaries.push_back(DirichletBC{node, 0.0, std::nullopt, number});
                } else if (kind == "roller_y" || kind == "symmetry_y") {
                    model.boundaries.push_back(DirichletBC{node, std::nullopt, 0.0, number});
                } else {
                    at.fail("boundary type must be fixed, pinned, roller_x, roller_y, symmetry_x, or symmetry_y");
                }
            } else {
                at.expect(t, 3, "node_id ux uy or node_id boundary_type");
                model.boundaries.push_back(DirichletBC{at.integer(t[0], "boundary node id"),
                                                       at.constraint(t[1], "ux"), at.constraint(t[2], "uy"),
                                                       number});
            }
            break;
        case Section::Loads:
            at.expect(t, 3, "node_id fx fy");
            model.loads.push_back(
                NodalLoad{at.integer(t[0], "load node id"), at.real(t[1], "fx"), at.real(t[2], "fy"), number});
            break;
        case Section::BodyForces:
            at.expect(t, 3, "element_id bx by");
            model.body_forces.push_back(
                BodyForce{at.integer(t[0], "element id"), at.real(t[1], "bx"), at.real(t[2], "by"), number});
            break;
        case Section::EdgeLoads: {
            if (t.size() < 4)
                at.fail("expected edge load: element_id edge traction tx ty or element_id edge pressure p");
            const int element = at.integer(t[0], "element id");
            const int edge = at.integer(t[1], "edge");
            const std::string kind = lowercase(t[2]);
            if (kind == "traction") {
                at.expect(t, 5, "element_id edge traction tx ty");
                model.edge_loads.push_back(EdgeLoad{element, edge, EdgeLoadType::Traction, at.real(t[3], "tx"),
                                                    at.real(t[4], "ty"), number});
            } else if (kind == "pressure") {
                at.expect(t, 4, "element_id edge pressure p");
                model.edge_loads.push_back(
                    EdgeLoad{element, edge, EdgeLoadType::Pressure, at.real(t[3], "pressure"), 0.0, number});
            } else {
                at.fail("edge load type must be traction or pressure");
            }
            break;
        }
        case Section::None:
            break;
    }
}


// ---- fast path for whole files (same grammar and values; any line it does not
// take verbatim sends the file through the line-by-line parser, which then reports
// the same first error the reference would)

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

struct Span {
    const char* b;
    const char* e;
};

int split_tokens(Span line, Span* out, int max_tokens) {
    int n = 0;
    const char* p = line.b;
    while (p < line.e) {
        while (p < line.e && is_space(*p)) ++p;
        if (p == line.e) break;
        const char* q = p;
        while (q < line.e && !is_space(*q)) ++q;
        if (n < max_tokens) out[n] = Span{p, q};
        ++n;
        p = q;
    }
    return n;
}

bool is_digit(char c) { return c >= '0' && c <= '9'; }

// std::stoll semantics on plain decimal tokens that fit an int.
bool fast_int(Span t, int& value) {
    const char* p = t.b;
    bool negative = false;
    if (p < t.e && (*p == '+' || *p == '-')) negative = *p++ == '-';
    if (p == t.e) return false;
    for (const char* q = p; q < t.e; ++q)
        if (!is_digit(*q)) return false;
    long long v = 0;
    const auto r = std::from_chars(p, t.e, v);
    if (r.ec != std::errc{} || r.ptr != t.e) return false;
    if (negative) v = -v;
    if (v < std::numeric_limits<int>::min() || v > std::numeric_limits<int>::max()) return false;
    value = static_cast<int>(v);
    return true;
}

// std::stod semantics on plain decimal tokens whose value is a finite normal double
// or an exact zero (both round correctly to nearest); anything else (hex, inf, nan,
// overflow, underflow, subnormal) goes to the line parser.
bool fast_real(Span t, double& value) {
    const char* p = t.b;
    if (p < t.e && *p == '+') ++p;
    const char* q = p;
    if (q < t.e && *q == '-') ++q;
    bool nonzero = false, digits = false;
    while (q < t.e && is_digit(*q)) {
        nonzero = nonzero || *q != '0';
        digits = true;
        ++q;
    }
    if (q < t.e && *q == '.') {
        ++q;
        while (q < t.e && is_digit(*q)) {
            nonzero = nonzero || *q != '0';
            digits = true;
            ++q;
        }
    }
    if (!digits) return false;
    if (q < t.e && (*q == 'e' || *q == 'E')) {
        ++q;
        if (q < t.e && (*q == '+' || *q == '-')) ++q;
        if (q == t.e || !is_digit(*q)) return false;
        while (q < t.e && is_digit(*q)) ++q;
    }
    if (q != t.e) return false;
    double v = 0.0;
    const auto r = std::from_chars(p, t.e, v, std::chars_format::general);
    if (r.ec != std::errc{} || r.ptr != t.e || !std::isfinite(v)) return false;
    if (v == 0.0 ? nonzero : std::abs(v) < std::numeric_limits<double>::min()) return false;
    value = v;
    return true;
}

struct Lines {
    std::vector<Span> text;
    std::vector<std::size_t> number;
};

bool fast_parse(const std::string& text, const std::string& source, Model& model) {
    Section section = Section::None;
    std::set<std::string> seen;
    Lines nodes, materials, elements;
    std::size_t number = 0;
    const char* p = text.data();
    const char* const end = p + text.size();
    try {
        while (p < end) {
            const char* eol = static_cast<const char*>(std::memchr(p, '\n', static_cast<std::size_t>(end - p)));
            if (!eol) eol = end;
            ++number;
            const char* e = static_cast<const char*>(std::memchr(p, '#', static_cast<std::size_t>(eol - p)));
            if (!e) e = eol;
            const char* b = p;
            p = eol < end ? eol + 1 : end;
            while (b < e && is_space(*b)) ++b;
            while (e > b && is_space(e[-1])) --e;
            if (b == e) continue;
            const LineReader at(source, number);
            if (*b == '[') {
                if (e[-1] != ']') return false;
                section = section_named(std::string(b + 1, e - 1), at);
                continue;
            }
            switch (section) {
                case Section::None: return false;
                case Section::Nodes:
                    nodes.text.push_back(Span{b, e});
                    nodes.number.push_back(number);
                    break;
                case Section::Materials:
                    materials.text.push_back(Span{b, e});
                    materials.number.push_back(number);
                    break;
                case Section::Elements:
                    elements.text.push_back(Span{b, e});
                    elements.number.push_back(number);
                    break;
                default: {
                    std::vector<std::string> t;
                    Span tok[16];
                    const int n = split_tokens(Span{b, e}, tok, 16);
                    if (n > 16) return false;
                    for (int k = 0; k < n; ++k) t.emplace_back(tok[k].b, tok[k].e);
                    apply_line(model, section, t, at, seen, number);
                }
            }
        }
    } catch (const ParseError&) {
        return false;
    }
    const std::size_t max = std::numeric_limits<std::size_t>::max();
    int failed = 0;
    model.nodes.resize(nodes.text.size());
#pragma omp parallel for schedule(static) reduction(| : failed)
    for (std::size_t i = 0; i < nodes.text.size(); ++i) {
        Span tok[3];
        Node& node = model.nodes[i];
        if (split_tokens(nodes.text[i], tok, 3) != 3 || !fast_int(tok[0], node.id) ||
            !fast_real(tok[1], node.position.x) || !fast_real(tok[2], node.position.y)) {
            failed |= 1;
            continue;
        }
        node.dof_x = max;
        node.dof_y = max;
        node.source_line = nodes.number[i];
    }
    model.materials.resize(materials.text.size());
#pragma omp parallel for schedule(static) reduction(| : failed)
    for (std::size_t i = 0; i < materials.text.size(); ++i) {
        Span tok[4];
        Material& m = model.materials[i];
        if (split_tokens(materials.text[i], tok, 4) != 4 || !fast_int(tok[0], m.id) ||
            !fast_real(tok[1], m.young_modulus) || !fast_real(tok[2], m.poisson_ratio) ||
            !fast_real(tok[3], m.thickness)) {
            failed |= 1;
            continue;
        }
        m.source_line = materials.number[i];
    }
    model.elements.resize(elements.text.size());
#pragma omp parallel for schedule(static) reduction(| : failed)
    for (std::size_t i = 0; i < elements.text.size(); ++i) {
        Span tok[6];
        Q4Element& el = model.elements[i];
        if (split_tokens(elements.text[i], tok, 6) != 6 || !fast_int(tok[0], el.id) ||
            !fast_int(tok[1], el.node_ids[0]) || !fast_int(tok[2], el.node_ids[1]) ||
            !fast_int(tok[3], el.node_ids[2]) || !fast_int(tok[4], el.node_ids[3]) ||
            !fast_int(tok[5], el.material_id)) {
            failed |= 1;
            continue;
        }
        el.source_line = elements.number[i];
    }
    return failed == 0;
}

}  // namespace

ParseError::ParseError(std::string source, std::size_t line, std::string message)
    : std::runtime_error(source + ":" + std::to_string(line) + ": " + message), source_(std::move(source)), line_(line) {}

Model InputParser::parse_file(const std::string& path) const {
    std::ifstream input(path, std::ios::binary);
    if (!input) throw ParseError(path, 0, "cannot open input file");
    std::string text;
    input.seekg(0, std::ios::end);
    const std::streamoff size = input.tellg();
    if (size > 0) {
        text.resize(static_cast<std::size_t>(size));
        input.seekg(0, std::ios::beg);
        input.read(text.data(), size);
        if (!input) throw ParseError(path, 0, "cannot open input file");
    }
    Model model;
    model.source_name = path;
    if (fast_parse(text, path, model)) return model;
    std::istringstream lines(text);
    return parse(lines, path);
}

Model InputParser::parse(std::istream& input, std::string source_name) const {
    Model model;
    model.source_name = source_name;
    Section section = Section::None;
    std::set<std::string> seen;
    std::string raw;
    std::size_t number = 0;
    while (std::getline(input, raw)) {
        ++number;
        const LineReader at(source_name, number);
        const std::size_t hash = raw.find('#');
        const std::string line = strip(hash == std::string::npos ? raw : raw.substr(0, hash));
        if (line.empty()) continue;
        if (line.front() == '[') {
            if (line.back() != ']') at.fail("section header must end with ']'");
            section = section_named(line.substr(1, line.size() - 2), at);
            continue;
        }
        if (section == Section::None) at.fail("data appears before the first section header");
        std::vector<std::string> t;
        {
            std::istringstream words(line);
            for (std::string w; words >> w;) t.push_back(w);
        }
        apply_line(model, section, t, at, seen, number);
    }
    return model;
}

}  // namespace lightfem
