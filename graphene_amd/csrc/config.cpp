// JSON-with-comments parser and the solver configuration tree (see config.hpp).
#include "config.hpp"

#include <cmath>
#include <cctype>
#include <cstdlib>

namespace graphene {

using gph::fail;
using gph::kInvalid;

namespace {
struct Parser {
  const std::string& s;
  size_t i = 0;
  explicit Parser(const std::string& t) : s(t) {}

  [[noreturn]] void error(const std::string& what) {
    fail(kInvalid, "JSON parse error at offset " + std::to_string(i) + ": " + what);
  }
  void skip() {
    for (;;) {
      while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' ||
                              s[i] == '\r'))
        i++;
      if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '/') {
        while (i < s.size() && s[i] != '\n') i++;
        continue;
      }
      if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '*') {
        size_t e = s.find("*/", i + 2);
        if (e == std::string::npos) error("unterminated comment");
        i = e + 2;
        continue;
      }
      break;
    }
  }
  Json value() {
    skip();
    if (i >= s.size()) error("unexpected end");
    char c = s[i];
    if (c == '{') return object();
    if (c == '[') return array();
    if (c == '"') {
      Json j;
      j.kind = Json::String;
      j.str = string();
      return j;
    }
    if (s.compare(i, 4, "true") == 0) {
      i += 4;
      Json j;
      j.kind = Json::Bool;
      j.b = true;
      return j;
    }
    if (s.compare(i, 5, "false") == 0) {
      i += 5;
      Json j;
      j.kind = Json::Bool;
      return j;
    }
    if (s.compare(i, 4, "null") == 0) {
      i += 4;
      return Json();
    }
    return number();
  }
  std::string string() {
    if (s[i] != '"') error("expected string");
    i++;
    std::string out;
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\') {
        i++;
        if (i >= s.size()) error("bad escape");
        char e = s[i];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (i + 4 >= s.size()) error("bad \\u escape");
            unsigned v = (unsigned)std::strtoul(s.substr(i + 1, 4).c_str(), nullptr, 16);
            if (v < 0x80) out += (char)v;
            else if (v < 0x800) { out += (char)(0xC0 | (v >> 6)); out += (char)(0x80 | (v & 0x3F)); }
            else { out += (char)(0xE0 | (v >> 12)); out += (char)(0x80 | ((v >> 6) & 0x3F)); out += (char)(0x80 | (v & 0x3F)); }
            i += 4;
            break;
          }
          default: out += e;
        }
        i++;
      } else {
        out += s[i++];
      }
    }
    if (i >= s.size()) error("unterminated string");
    i++;
    return out;
  }
  Json number() {
    size_t st = i;
    if (i < s.size() && (s[i] == '-' || s[i] == '+')) i++;
    while (i < s.size() && (std::isdigit((unsigned char)s[i]) || s[i] == '.' ||
                            s[i] == 'e' || s[i] == 'E' || s[i] == '-' || s[i] == '+'))
      i++;
    if (st == i) error("unexpected character");
    Json j;
    j.kind = Json::Number;
    j.str = s.substr(st, i - st);
    char* end = nullptr;
    j.num = std::strtod(j.str.c_str(), &end);
    if (!end || *end != '\0') error("bad number " + j.str);
    return j;
  }
  Json array() {
    Json j;
    j.kind = Json::Array;
    i++;
    skip();
    if (i < s.size() && s[i] == ']') {
      i++;
      return j;
    }
    for (;;) {
      j.arr.push_back(value());
      skip();
      if (i < s.size() && s[i] == ',') {
        i++;
        skip();
        if (i < s.size() && s[i] == ']') {  // trailing comma
          i++;
          return j;
        }
        continue;
      }
      if (i < s.size() && s[i] == ']') {
        i++;
        return j;
      }
      error("expected , or ]");
    }
  }
  Json object() {
    Json j;
    j.kind = Json::Object;
    i++;
    skip();
    if (i < s.size() && s[i] == '}') {
      i++;
      return j;
    }
    for (;;) {
      skip();
      std::string k = string();
      skip();
      if (i >= s.size() || s[i] != ':') error("expected :");
      i++;
      j.obj[k] = value();
      skip();
      if (i < s.size() && s[i] == ',') {
        i++;
        skip();
        if (i < s.size() && s[i] == '}') {
          i++;
          return j;
        }
        continue;
      }
      if (i < s.size() && s[i] == '}') {
        i++;
        return j;
      }
      error("expected , or }");
    }
  }
};

// Field readers with boost::property_tree::get_optional semantics
// (lg/matrix/solver/Configuration.cpp:37-101): absent or unconvertible
// values leave the default in place.
void readFloat(const Json& c, const char* k, float& v) {
  const Json* j = c.get(k);
  if (!j) return;
  if (j->kind == Json::Number) v = (float)j->num;
  else if (j->kind == Json::String) {
    char* e = nullptr;
    double d = std::strtod(j->str.c_str(), &e);
    if (e && *e == '\0' && !j->str.empty()) v = (float)d;
  }
}
void readInt(const Json& c, const char* k, int& v) {
  const Json* j = c.get(k);
  if (!j) return;
  if (j->kind == Json::Number) {
    // ptree parses the literal text as an integer; "1e3" or "4.5" fail
    const std::string& t = j->str;
    char* e = nullptr;
    long long x = std::strtoll(t.c_str(), &e, 10);
    if (e && *e == '\0' && x >= INT_MIN && x <= INT_MAX) v = (int)x;
  }
}
void readBool(const Json& c, const char* k, bool& v) {
  const Json* j = c.get(k);
  if (!j) return;
  if (j->kind == Json::Bool) v = j->b;
  else if (j->kind == Json::String) {
    if (j->str == "true" || j->str == "1") v = true;
    if (j->str == "false" || j->str == "0") v = false;
  } else if (j->kind == Json::Number && (j->str == "0" || j->str == "1")) {
    v = j->str == "1";
  }
}
std::string asText(const Json& j) {
  if (j.kind == Json::String) return j.str;
  if (j.kind == Json::Bool) return j.b ? "true" : "false";
  if (j.kind == Json::Number) return j.str;
  return "";
}
void readNorm(const Json& c, const char* k, VectorNorm& v) {
  const Json* j = c.get(k);
  if (j) v = parseVectorNorm(asText(*j));
}
// twofloat32 is accepted as an extended precision only (the reference's
// inferType admits it with float32 operands only, dsl/common/details/
// Expressions.cpp:41-52; the working precision is that of x)
void readType(const Json& c, const char* k, std::optional<gph::DType>& v,
              bool extended = false) {
  const Json* j = c.get(k);
  if (!j) return;
  v = parseType(asText(*j));
  if (*v == gph::DType::TF32 && !extended)
    fail(gph::kUnsupported, std::string(k) + ": twofloat32 is supported as extendedPrecisionType only");
}
// Configuration.cpp:51-77
void readMulticolor(const Json& c, const char* k,
                    matrix::solver::MultiColorMode& v) {
  using M = matrix::solver::MultiColorMode;
  const Json* j = c.get(k);
  if (!j) return;
  std::string s = asText(*j);
  if (s == "true" || s == "on") v = M::On;
  else if (s == "false" || s == "off") v = M::Off;
  else if (s == "auto") v = M::Auto;
  else
    fail(kInvalid, std::string("Invalid value for MultiColorMode field ") + k +
                       ": " + s + ". Allowed values are true, false, on, off or \"auto\"");
}
}  // namespace

Json Json::parse(const std::string& text) {
  Parser p(text);
  Json j = p.value();
  p.skip();
  if (p.i != text.size()) p.error("trailing characters");
  return j;
}

VectorNorm parseVectorNorm(const std::string& s) {
  if (s == "L1") return VectorNorm::L1;
  if (s == "LINF") return VectorNorm::LINF;
  if (s == "L2") return VectorNorm::L2;
  fail(kInvalid, "Invalid norm: " + s);
}
std::string normToString(VectorNorm n) {
  switch (n) {
    case VectorNorm::L1: return "L1";
    case VectorNorm::L2: return "L2";
    case VectorNorm::LINF: return "LINF";
    case VectorNorm::None: return "None";
  }
  return "None";
}

// Floating-point names of lg/common/Type.cpp:69-88.
gph::DType parseType(const std::string& s) {
  if (s == "float16") return gph::DType::F16;
  if (s == "float32") return gph::DType::F32;
  if (s == "float64") return gph::DType::F64;
  if (s == "twofloat32") return gph::DType::TF32;
  fail(kInvalid, "Unknown type in config: " + s);
}

namespace matrix::solver {

pbicgstab::Configuration::Configuration(const Json& c) {
  readFloat(c, "absTolerance", absTolerance);
  readFloat(c, "relTolerance", relTolerance);
  readFloat(c, "relResidual", relResidual);
  readInt(c, "maxIterations", maxIterations);
  readInt(c, "minIterations", minIterations);
  readBool(c, "printPerformanceAfterSolve", printPerformanceAfterSolve);
  readBool(c, "printPerformanceEachIteration", printPerformanceEachIteration);
  readNorm(c, "norm", norm);
  readBool(c, "verbose", verbose);
  readType(c, "workingType", workingType);
  readType(c, "matrixValueType", matrixValueType);
  readBool(c, "relToleranceFromB", relToleranceFromB);
  readInt(c, "iterationsPerSync", iterationsPerSync);
  if (const Json* p = c.get("preconditioner"))
    preconditioner = solver::Configuration::fromPTree(*p);
}

iterativerefinement::Configuration::Configuration(const Json& c) {
  readFloat(c, "absTolerance", absTolerance);
  readFloat(c, "relTolerance", relTolerance);
  readFloat(c, "relResidual", relResidual);
  readInt(c, "maxIterations", maxIterations);
  readInt(c, "minIterations", minIterations);
  readBool(c, "printPerformanceAfterSolve", printPerformanceAfterSolve);
  readBool(c, "printPerformanceEachIteration", printPerformanceEachIteration);
  readNorm(c, "norm", norm);
  readType(c, "extendedPrecisionType", extendedPrecisionType, true);
  readType(c, "workingType", workingType);
  readType(c, "matrixValueType", matrixValueType);
  const Json* in = c.get("innerSolver");
  if (!in) fail(kInvalid, "IterativeRefinement requires an innerSolver configuration");
  innerSolver = solver::Configuration::fromPTree(*in);
}

ilu::Configuration::Configuration(const Json& c) {
  readMulticolor(c, "solveMulticolor", solveMulticolor);
  readMulticolor(c, "factorizeMulticolor", factorizeMulticolor);
  readBool(c, "diagonalBased", diagonalBased);
  readType(c, "factorType", factorType);
  readInt(c, "rowsPerWave", rowsPerWave);

  if (const Json* j = c.get("substitution")) {
    substitution = asText(*j);
    if (substitution != "level" && substitution != "packed" && substitution != "tile" &&
        substitution != "rows" && substitution != "block")
      fail(kInvalid, "Invalid value for substitution: " + substitution +
                         ". Allowed values are \"packed\", \"rows\", \"level\", \"block\" "
                         "or \"tile\"");
  }
}

gaussseidel::Configuration::Configuration(const Json& c) {
  readFloat(c, "absTolerance", absTolerance);
  readFloat(c, "relTolerance", relTolerance);
  readFloat(c, "relResidual", relResidual);
  readInt(c, "maxIterations", maxIterations);
  readInt(c, "minIterations", minIterations);
  readInt(c, "numSweeps", numSweeps);
  readInt(c, "numFixedIterations", numFixedIterations);
  readMulticolor(c, "solveMulticolor", solveMulticolor);
  readBool(c, "printPerformanceAfterSolve", printPerformanceAfterSolve);
  readBool(c, "printPerformanceEachIteration", printPerformanceEachIteration);
  readNorm(c, "norm", norm);
  readType(c, "workingType", workingType);
  readBool(c, "symmetric", symmetric);
}

restarter::Configuration::Configuration(const Json& c) {
  readInt(c, "maxRestarts", maxRestarts);
  const Json* in = c.get("innerSolver");
  if (!in) fail(kInvalid, "Restarter requires an innerSolver configuration");
  innerSolver = solver::Configuration::fromPTree(*in);
}

// Configuration.cpp:103-124
std::shared_ptr<Configuration> Configuration::fromPTree(const Json& c) {
  const Json* t = c.get("type");
  if (!t) fail(kInvalid, "No solver type specified in configuration");
  std::string solver = asText(*t);
  if (solver == "IterativeRefinement")
    return std::make_shared<iterativerefinement::Configuration>(c);
  if (solver == "GaussSeidel")
    return std::make_shared<gaussseidel::Configuration>(c);
  if (solver == "ILU") return std::make_shared<ilu::Configuration>(c);
  if (solver == "PBiCGStab")
    return std::make_shared<pbicgstab::Configuration>(c);
  if (solver == "restarter")
    return std::make_shared<restarter::Configuration>(c);
  fail(kInvalid, "Unknown solver: " + solver);
}

std::shared_ptr<Configuration> Configuration::fromJSON(const std::string& text) {
  return fromPTree(Json::parse(text));
}

}  // namespace matrix::solver
}  // namespace graphene
