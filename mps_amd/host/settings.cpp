#include "settings.hpp"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "logger.hpp"

Domain::Domain(const double& xMin, const double& xMax, const double& yMin, const double& yMax, const double& zMin,
               const double& zMax) {
    x.min = xMin;
    x.max = xMax;
    x.length = xMax - xMin;
    y.min = yMin;
    y.max = yMax;
    y.length = yMax - yMin;
    z.min = zMin;
    z.max = zMax;
    z.length = zMax - zMin;
}

namespace {

// Block mappings by indentation, "key: value" with keys that may hold spaces,
// flow sequences "[a, b, c]", "#" comments, plain scalars.  That is all the
// reference's settings files use (tutorials/*/settings.yml).
struct Node {
    std::string scalar;
    std::vector<std::string> seq;
    std::map<std::string, std::shared_ptr<Node>> map;
    bool is_seq = false;
};

struct TypeError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char) s[a])) a++;
    while (b > a && std::isspace((unsigned char) s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::string strip_comment(const std::string& s) {
    bool sq = false, dq = false;
    for (size_t k = 0; k < s.size(); k++) {
        const char c = s[k];
        if (c == '\'' && !dq) sq = !sq;
        if (c == '"' && !sq) dq = !dq;
        if (c == '#' && !sq && !dq && (k == 0 || std::isspace((unsigned char) s[k - 1]))) return s.substr(0, k);
    }
    return s;
}

std::string unquote(const std::string& s) {
    if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
        return s.substr(1, s.size() - 2);
    return s;
}

std::shared_ptr<Node> parse(std::istream& in) {
    auto root = std::make_shared<Node>();
    std::vector<std::pair<int, std::shared_ptr<Node>>> stack{{-1, root}};
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        lineno++;
        const std::string body = strip_comment(line);
        if (trim(body).empty()) continue;
        int indent = 0;
        while (indent < (int) body.size() && body[indent] == ' ') indent++;
        const std::string t = trim(body);
        const size_t colon = t.find(':');
        if (colon == std::string::npos) throw std::runtime_error("line " + std::to_string(lineno) + ": expected key: value");
        const std::string key = trim(t.substr(0, colon));
        const std::string val = trim(t.substr(colon + 1));
        while (stack.size() > 1 && stack.back().first >= indent) stack.pop_back();
        auto parent = stack.back().second;
        auto node = std::make_shared<Node>();
        if (!val.empty() && val.front() == '[') {
            if (val.back() != ']') throw std::runtime_error("line " + std::to_string(lineno) + ": unterminated sequence");
            node->is_seq = true;
            std::stringstream ss(val.substr(1, val.size() - 2));
            std::string item;
            while (std::getline(ss, item, ',')) node->seq.push_back(unquote(trim(item)));
        } else {
            node->scalar = unquote(val);
        }
        parent->map[key] = node;
        if (val.empty()) stack.push_back({indent, node});
    }
    return root;
}

const Node& at(const Node& n, const std::string& key) {
    auto it = n.map.find(key);
    if (it == n.map.end()) throw std::runtime_error("key not found: " + key);
    return *it->second;
}

// fkYAML types a plain scalar: a double needs a decimal point or an exponent
// (the reference's settings files note this, tutorials/dambreak/settings.yml:1-3).
double as_double(const std::string& s, const std::string& what) {
    const bool looks_float = s.find('.') != std::string::npos || s.find('e') != std::string::npos ||
                             s.find('E') != std::string::npos || s == ".inf" || s == "-.inf" || s == ".nan";
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (!looks_float || end == s.c_str() || *end != '\0') throw TypeError("type_error: " + what + " is not a float: " + s);
    return v;
}

int as_int(const std::string& s, const std::string& what) {
    char* end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end != '\0') throw TypeError("type_error: " + what + " is not an integer: " + s);
    return (int) v;
}

bool as_bool(const std::string& s, const std::string& what) {
    if (s == "true" || s == "True" || s == "TRUE") return true;
    if (s == "false" || s == "False" || s == "FALSE") return false;
    throw TypeError("type_error: " + what + " is not a boolean: " + s);
}

double seq_double(const Node& n, size_t k, const std::string& what) {
    if (!n.is_seq || k >= n.seq.size()) throw std::runtime_error("out_of_range: " + what);
    return as_double(n.seq[k], what);
}

}  // namespace

bool Settings::load(const std::filesystem::path& inputYamlPath) {
    std::ifstream ifs(inputYamlPath);
    if (!ifs.is_open()) mpslog::error("Could not open setting file: " + inputYamlPath.string());
    try {
        auto rootp = parse(ifs);
        const Node& root = *rootp;
        dim = as_int(at(root, "dim").scalar, "dim");
        particleDistance = as_double(at(root, "particle distance").scalar, "particle distance");
        dt = as_double(at(root, "dt").scalar, "dt");
        finishTime = as_double(at(root, "finish time").scalar, "finish time");
        outputInterval = as_double(at(root, "output interval").scalar, "output interval");
        cflCondition = as_double(at(root, "CFL condition").scalar, "CFL condition");

        const double xMin = seq_double(at(root, "domain range X"), 0, "domain range X");
        const double xMax = seq_double(at(root, "domain range X"), 1, "domain range X");
        const double yMin = seq_double(at(root, "domain range Y"), 0, "domain range Y");
        const double yMax = seq_double(at(root, "domain range Y"), 1, "domain range Y");
        const double zMin = seq_double(at(root, "domain range Z"), 0, "domain range Z");
        const double zMax = seq_double(at(root, "domain range Z"), 1, "domain range Z");
        domain = Domain(xMin, xMax, yMin, yMax, zMin, zMax);

        density = as_double(at(root, "density").scalar, "density");
        kinematicViscosity = as_double(at(root, "kinematic viscosity").scalar, "kinematic viscosity");

        gravity.x() = seq_double(at(root, "gravity"), 0, "gravity");
        gravity.y() = seq_double(at(root, "gravity"), 1, "gravity");
        gravity.z() = seq_double(at(root, "gravity"), 2, "gravity");

        const Node& er = at(root, "effective radius ratio");
        effectiveRadius.pressure = as_double(at(er, "pressure").scalar, "pressure") * particleDistance;
        effectiveRadius.viscosity = as_double(at(er, "viscosity").scalar, "viscosity") * particleDistance;
        effectiveRadius.surfaceDetection =
            as_double(at(er, "surface detection").scalar, "surface detection") * particleDistance;
        effectiveRadius.max =
            std::max({effectiveRadius.pressure, effectiveRadius.viscosity, effectiveRadius.surfaceDetection});

        thresholdForSurfaceDetection =
            as_double(at(root, "threshold for surface detection").scalar, "threshold for surface detection");

        const Node& ps = at(root, "pressure stabilization");
        const Node& qc = at(ps, "quasi compressibility");
        quasiCompressibility.on = as_bool(at(qc, "on").scalar, "on");
        if (quasiCompressibility.on)
            quasiCompressibility.compressibility = as_double(at(qc, "compressibility").scalar, "compressibility");
        const Node& hos = at(ps, "higher order source term");
        higherOrderSourceTerm.on = as_bool(at(hos, "on").scalar, "on");
        if (higherOrderSourceTerm.on) higherOrderSourceTerm.gamma = as_double(at(hos, "gamma").scalar, "gamma");
        relaxationCoefficientForPressure =
            as_double(at(root, "relaxation coefficient for pressure").scalar, "relaxation coefficient for pressure");

        collisionDistance = as_double(at(root, "collision distance ratio").scalar, "collision distance ratio") *
                            particleDistance;
        coefficientOfRestitution =
            as_double(at(root, "coefficient of restitution").scalar, "coefficient of restitution");

        inputCsvPath = at(root, "inputCsvPath").scalar;
    } catch (const std::exception& e) {
        mpslog::error(e.what());
        return false;
    }
    return true;
}

mps_settings Settings::toC() const {
    mps_settings s{};
    s.dim = (int32_t) dim;
    s.qc_on = quasiCompressibility.on ? 1 : 0;
    s.hos_on = higherOrderSourceTerm.on ? 1 : 0;
    s.particle_distance = particleDistance;
    s.dt = dt;
    s.finish_time = finishTime;
    s.output_interval = outputInterval;
    s.cfl_condition = cflCondition;
    s.domain_min[0] = domain.x.min;
    s.domain_min[1] = domain.y.min;
    s.domain_min[2] = domain.z.min;
    s.domain_max[0] = domain.x.max;
    s.domain_max[1] = domain.y.max;
    s.domain_max[2] = domain.z.max;
    s.density = density;
    s.kinematic_viscosity = kinematicViscosity;
    for (int k = 0; k < 3; k++) s.gravity[k] = gravity[k];
    s.re_pressure = effectiveRadius.pressure;
    s.re_viscosity = effectiveRadius.viscosity;
    s.re_surface = effectiveRadius.surfaceDetection;
    s.re_max = effectiveRadius.max;
    s.surface_threshold = thresholdForSurfaceDetection;
    s.compressibility = quasiCompressibility.on ? quasiCompressibility.compressibility : 0.0;
    s.hos_gamma = higherOrderSourceTerm.on ? higherOrderSourceTerm.gamma : 0.0;
    s.relaxation = relaxationCoefficientForPressure;
    s.collision_distance = collisionDistance;
    s.restitution = coefficientOfRestitution;
    return s;
}
