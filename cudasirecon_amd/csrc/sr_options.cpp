// cudasirecon option parser (see sr_options.hpp).
#include "sr_options.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>

#include "errors.hpp"

namespace simr {
namespace cli {

namespace {

enum class Kind {
  Str,           // value<std::string>
  Int,           // value<int>
  Float,         // value<float>
  Short,         // value<short>
  Unsigned,      // value<unsigned>
  Switch,        // bool_switch: no value on the command line, bool in a config file
  IntImplicit0,  // value<int>()->implicit_value(false): a bare flag stores 0
  BoolImplicit0, // value<bool>()->implicit_value(false)
  FloatList,     // value<std::vector<float>>()->multitoken()
  Flag           // help / version: no value
};

struct Opt {
  const char* name;
  char shortc;
  Kind kind;
  const char* deflt;  // nullptr: no default (the bound field keeps its value)
  bool required;
  const char* help;
};

// The option table of setupProgramOptions (cudaSirecon.cpp:913-1009), in its order.
const Opt kOpts[] = {
    {"input-file", 0, Kind::Str, nullptr, false, "input file name (or data folder in TIFF mode)"},
    {"output-file", 0, Kind::Str, nullptr, false, "output file name (or file name pattern in TIFF mode)"},
    {"otf-file", 0, Kind::Str, nullptr, true, "OTF file name"},
    {"config", 'c', Kind::Str, "", false, "config file of key=value lines, in place of command-line options"},
    {"ndirs", 0, Kind::Int, "3", false, "number of SIM directions"},
    {"nphases", 0, Kind::Int, "5", false, "number of pattern phases per direction"},
    {"nordersout", 0, Kind::Int, "0", false, "number of output SIM orders (0: nphases/2+1)"},
    {"angle0", 0, Kind::Float, "1.648", false, "angle of the first SIM direction in radians"},
    {"ls", 0, Kind::Float, "0.172", false, "line spacing of the SIM pattern in microns"},
    {"na", 0, Kind::Float, "1.2", false, "detection numerical aperture"},
    {"nimm", 0, Kind::Float, "1.33", false, "refractive index of the immersion medium"},
    {"wiener", 0, Kind::Float, "0.01", false, "Wiener constant"},
    {"otfcutoff", 0, Kind::Float, "0.006", false, "OTF threshold below which makeoverlaps ignores a frequency"},
    {"zoomfact", 0, Kind::Float, "2", false, "lateral zoom of the output over the input"},
    {"zzoom", 0, Kind::Int, "1", false, "axial zoom of the output"},
    {"background", 0, Kind::Float, "0", false, "camera background"},
    {"usecorr", 0, Kind::Str, nullptr, false, "flat-field correction file (MRC: background, slope)"},
    {"forcemodamp", 0, Kind::FloatList, nullptr, false, "force the modulation amplitudes of orders 1.. to at least these"},
    {"k0angles", 0, Kind::Str, nullptr, false, "comma-separated k0 angles of all directions (radians)"},
    {"otfRA", 0, Kind::Switch, "0", false, "use the rotationally averaged OTF"},
    {"otfPerAngle", 0, Kind::Switch, "0", false, "one OTF per SIM direction"},
    {"fastSI", 0, Kind::Switch, "0", false, "raw images in Z->Angle->Phase order (else Angle->Z->Phase)"},
    {"k0searchAll", 0, Kind::IntImplicit0, nullptr, false, "search for k0 at every time point"},
    {"norescale", 0, Kind::IntImplicit0, nullptr, false, "no bleach correction"},
    {"equalizez", 0, Kind::Switch, "0", false, "bleach correction across z"},
    {"equalizet", 0, Kind::Switch, "0", false, "bleach correction across time"},
    {"dampenOrder0", 0, Kind::Switch, "0", false, "dampen order 0 in the final assembly"},
    {"nosuppress", 0, Kind::IntImplicit0, nullptr, false, "do not suppress the DC singularity of the bands"},
    {"nokz0", 0, Kind::Switch, "0", false, "drop the kz=0 plane of order 0"},
    {"gammaApo", 0, Kind::Float, "1", false, "output apodization gamma"},
    {"explodefact", 0, Kind::Float, "1", false, "scale the reciprocal-space distance between orders (debug)"},
    {"nofilterovlps", 0, Kind::BoolImplicit0, nullptr, false, "do not filter the overlap of bands (debug)"},
    {"saveprefiltered", 0, Kind::Str, nullptr, false, "save separated bands and exit (debug)"},
    {"savealignedraw", 0, Kind::Str, nullptr, false, "save drift-fixed raw data and exit (debug)"},
    {"saveoverlaps", 0, Kind::Str, nullptr, false, "save overlaps and exit (debug)"},
    {"2lenses", 0, Kind::Switch, "0", false, "I5S data"},
    {"bessel", 0, Kind::Switch, "0", false, "Bessel-SIM data"},
    {"besselExWave", 0, Kind::Float, "0.488", false, "Bessel excitation wavelength in microns"},
    {"besselNA", 0, Kind::Float, "0.144", false, "Bessel excitation NA"},
    {"deskew", 0, Kind::Float, "0", false, "deskew angle (0: no deskew)"},
    {"deskewshift", 0, Kind::Int, "0", false, "extra x shift of the deskewed output"},
    {"noRecon", 0, Kind::Switch, "0", false, "no reconstruction (with --deskew)"},
    {"cropXY", 0, Kind::Unsigned, "0", false, "crop x-y to this size (0: no cropping)"},
    {"xyres", 0, Kind::Float, "0.1", false, "x-y pixel size (TIFF input)"},
    {"zres", 0, Kind::Float, "0.2", false, "z pixel size (TIFF input)"},
    {"zresPSF", 0, Kind::Float, "0.15", false, "z pixel size of the PSF (TIFF OTF)"},
    {"wavelength", 0, Kind::Short, "530", false, "emission wavelength in nm (TIFF input)"},
    {"writeTitle", 0, Kind::Switch, "0", false, "write the command line into the MRC header"},
    {"help", 'h', Kind::Flag, nullptr, false, "produce this help message"},
    {"version", 0, Kind::Flag, nullptr, false, "show the version"},
};

const char* kPositional[] = {"input-file", "output-file", "otf-file"};

const Opt* find_exact(const std::string& name) {
  for (const Opt& o : kOpts)
    if (name == o.name) return &o;
  return nullptr;
}

const Opt* find_short(char c) {
  for (const Opt& o : kOpts)
    if (o.shortc == c) return &o;
  return nullptr;
}

// Long-option lookup with unambiguous-prefix guessing, as the default
// command-line style allows.
const Opt* find_long(const std::string& name) {
  if (const Opt* o = find_exact(name)) return o;
  std::vector<const Opt*> hits;
  for (const Opt& o : kOpts)
    if (std::strncmp(o.name, name.c_str(), name.size()) == 0) hits.push_back(&o);
  if (hits.size() == 1) return hits[0];
  if (hits.empty()) throw ArgError("unrecognised option '--" + name + "'");
  std::string m = "option '--" + name + "' is ambiguous and matches ";
  for (size_t i = 0; i < hits.size(); i++) m += (i ? ", '--" : "'--") + std::string(hits[i]->name) + "'";
  throw ArgError(m);
}

bool takes_value(Kind k) { return k != Kind::Switch && k != Kind::Flag; }

bool looks_numeric(const std::string& s) {
  if (s.size() < 2 || s[0] != '-') return false;
  char* end = nullptr;
  errno = 0;
  std::strtod(s.c_str(), &end);
  return end && *end == '\0';
}

struct Stored {
  std::vector<std::string> values;
  int source = 0;  // 0 command line, 1 config file
};

using Store = std::map<std::string, Stored>;

std::string display(const Opt& o, int source) { return source == 0 ? "--" + std::string(o.name) : o.name; }

void add(Store& st, const Opt& o, const std::string& v, int source) {
  auto it = st.find(o.name);
  if (it != st.end()) {
    if (it->second.source != source) return;  // an earlier source wins
    if (o.kind != Kind::FloatList)
      throw ArgError("option '" + display(o, source) + "' cannot be specified more than once");
    it->second.values.push_back(v);
    return;
  }
  Stored s;
  s.values.push_back(v);
  s.source = source;
  st[o.name] = s;
}

[[noreturn]] void invalid(const Opt& o, const std::string& v, int source) {
  throw ArgError("the argument ('" + v + "') for option '" + display(o, source) + "' is invalid");
}

long long to_int(const Opt& o, const std::string& v, int source) {
  if (v.empty()) invalid(o, v, source);
  char* end = nullptr;
  errno = 0;
  const long long x = std::strtoll(v.c_str(), &end, 10);
  if (errno || !end || *end != '\0' || std::isspace((unsigned char)v[0])) invalid(o, v, source);
  return x;
}

float to_float(const Opt& o, const std::string& v, int source) {
  if (v.empty() || std::isspace((unsigned char)v[0])) invalid(o, v, source);
  char* end = nullptr;
  errno = 0;
  const double x = std::strtod(v.c_str(), &end);
  if (!end || *end != '\0') invalid(o, v, source);
  return (float)x;
}

bool to_bool(const Opt& o, const std::string& v, int source) {
  std::string s = v;
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  if (s.empty() || s == "on" || s == "yes" || s == "1" || s == "true") return true;
  if (s == "off" || s == "no" || s == "0" || s == "false") return false;
  invalid(o, v, source);
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}

// Command line: long (--name, --name=value, --name value), short (-c value,
// -cvalue) and positional arguments.
void parse_cli_tokens(int argc, char** argv, Store& st) {
  size_t npos = 0;
  for (int i = 1; i < argc; i++) {
    const std::string tok = argv[i] ? argv[i] : "";
    if (tok.size() > 2 && tok[0] == '-' && tok[1] == '-') {
      const size_t eq = tok.find('=');
      const std::string name = tok.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
      const Opt& o = *find_long(name);
      if (!takes_value(o.kind)) {
        if (eq != std::string::npos) throw ArgError("option '--" + std::string(o.name) + "' does not take any arguments");
        add(st, o, "", 0);
      } else if (eq != std::string::npos) {
        add(st, o, tok.substr(eq + 1), 0);
      } else if (o.kind == Kind::IntImplicit0 || o.kind == Kind::BoolImplicit0) {
        add(st, o, "0", 0);
      } else if (o.kind == Kind::FloatList) {
        int taken = 0;
        while (i + 1 < argc) {
          const std::string nx = argv[i + 1];
          char* end = nullptr;
          std::strtod(nx.c_str(), &end);
          if (nx.empty() || !end || *end != '\0') break;
          add(st, o, nx, 0);
          i++;
          taken++;
        }
        if (!taken) throw ArgError("the required argument for option '--" + std::string(o.name) + "' is missing");
      } else {
        if (i + 1 >= argc) throw ArgError("the required argument for option '--" + std::string(o.name) + "' is missing");
        const std::string nx = argv[i + 1];
        if (nx.size() > 1 && nx[0] == '-' && !looks_numeric(nx))
          throw ArgError("the required argument for option '--" + std::string(o.name) + "' is missing");
        add(st, o, nx, 0);
        i++;
      }
    } else if (tok.size() >= 2 && tok[0] == '-' && !looks_numeric(tok)) {
      const Opt* o = find_short(tok[1]);
      if (!o) throw ArgError("unrecognised option '" + tok.substr(0, 2) + "'");
      if (!takes_value(o->kind)) {
        if (tok.size() > 2) throw ArgError("unrecognised option '" + tok + "'");
        add(st, *o, "", 0);
      } else if (tok.size() > 2) {
        add(st, *o, tok.substr(2), 0);
      } else {
        if (i + 1 >= argc) throw ArgError("the required argument for option '--" + std::string(o->name) + "' is missing");
        add(st, *o, argv[++i], 0);
      }
    } else {
      if (npos >= 3) throw ArgError("too many positional options have been specified on the command line");
      add(st, *find_exact(kPositional[npos++]), tok, 0);
    }
  }
}

// Config file: key=value lines, '#' comments, [section] prefixes, exact names.
void parse_config_stream(std::istream& in, Store& st) {
  std::string line, prefix;
  while (std::getline(in, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.erase(hash);
    line = trim(line);
    if (line.empty()) continue;
    if (line.front() == '[' && line.back() == ']') {
      prefix = trim(line.substr(1, line.size() - 2));
      if (!prefix.empty()) prefix += '.';
      continue;
    }
    const size_t eq = line.find('=');
    if (eq == std::string::npos || trim(line.substr(0, eq)).empty())
      throw ArgError("the options configuration file contains an invalid line '" + line + "'");
    const std::string name = prefix + trim(line.substr(0, eq));
    const std::string value = trim(line.substr(eq + 1));
    const Opt* o = find_exact(name);
    if (!o || o->kind == Kind::Flag) throw ArgError("unrecognised option '" + name + "'");
    add(st, *o, value, 1);
  }
}

// notify(): write stored values (or defaults) into the bound fields.
void notify(const Store& st, SR_ReconParams& r, SR_ImageParams& img, std::string* config_file) {
  for (const Opt& o : kOpts) {
    auto it = st.find(o.name);
    const bool given = it != st.end();
    if (!given && !o.deflt) continue;
    const std::string v = given ? it->second.values.back() : o.deflt;
    const int src = given ? it->second.source : 0;
    const std::string n = o.name;
    auto I = [&]() { return (int)to_int(o, v, src); };
    auto F = [&]() { return to_float(o, v, src); };
    auto B = [&]() { return given ? (o.kind == Kind::Switch && src == 0 ? true : to_bool(o, v, src)) : false; };
    if (n == "config") {
      if (config_file) *config_file = v;
    } else if (n == "ndirs") r.ndirs = I();
    else if (n == "nphases") r.nphases = I();
    else if (n == "nordersout") r.norders_output = I();
    else if (n == "angle0") r.k0startangle = F();
    else if (n == "ls") r.linespacing = F();
    else if (n == "na") r.na = F();
    else if (n == "nimm") r.nimm = F();
    else if (n == "wiener") r.wiener = F();
    else if (n == "otfcutoff") r.otfcutoff = F();
    else if (n == "zoomfact") r.zoomfact = F();
    else if (n == "zzoom") r.z_zoom = I();
    else if (n == "background") r.constbkgd = F();
    else if (n == "otfRA") r.bRadAvgOTF = B();
    else if (n == "otfPerAngle") r.bOneOTFperAngle = B();
    else if (n == "fastSI") r.bFastSIM = B();
    else if (n == "k0searchAll") r.bUseTime0k0 = I();
    else if (n == "norescale") r.do_rescale = I();
    else if (n == "equalizez") r.equalizez = B();
    else if (n == "equalizet") r.equalizet = B();
    else if (n == "dampenOrder0") r.bDampenOrder0 = B();
    else if (n == "nosuppress") r.bSuppress_singularities = I();
    else if (n == "nokz0") r.bNoKz0 = B();
    else if (n == "gammaApo") r.apoGamma = F();
    else if (n == "explodefact") r.explodefact = F();
    else if (n == "nofilterovlps") r.bFilteroverlaps = to_bool(o, v, src);
    else if (n == "2lenses") r.bTwolens = B();
    else if (n == "bessel") r.bBessel = B();
    else if (n == "besselExWave") r.BesselLambdaEx = F();
    else if (n == "besselNA") r.BesselNA = F();
    else if (n == "deskew") r.deskewAngle = F();
    else if (n == "deskewshift") r.extraShift = I();
    else if (n == "noRecon") r.bNoRecon = B();
    else if (n == "cropXY") r.cropXYto = (uint32_t)to_int(o, v, src);
    else if (n == "xyres") img.dxy = F();
    else if (n == "zres") img.dz = F();
    else if (n == "zresPSF") r.dzPSF = F();
    else if (n == "wavelength") {
      const long long w = to_int(o, v, src);
      if (w < -32768 || w > 32767) invalid(o, v, src);
      img.wave[0] = (int16_t)w;
    } else if (n == "writeTitle") r.bWriteTitle = B();
  }
}

// setParams (cudaSirecon.cpp:1013-1071): the values without a bound field.
void set_params(const Store& st, SR_ReconParams& r, bool verbose) {
  auto get = [&](const char* name) -> const Stored* {
    auto it = st.find(name);
    return it == st.end() ? nullptr : &it->second;
  };
  if (auto s = get("input-file")) set_str(r.ifiles, s->values.back());
  if (auto s = get("output-file")) set_str(r.ofiles, s->values.back());
  if (auto s = get("otf-file")) set_str(r.otffiles, s->values.back());
  if (auto s = get("usecorr")) {
    set_str(r.corrfiles, s->values.back());
    r.bUsecorr = 1;
  }
  if (auto s = get("forcemodamp")) {
    const Opt& o = *find_exact("forcemodamp");
    if (s->values.size() > SR_MAX_LIST) throw ArgError("too many values for option '--forcemodamp'");
    r.nforceamp = (int)s->values.size();
    for (size_t i = 0; i < s->values.size(); i++) r.forceamp[i] = to_float(o, s->values[i], s->source);
  }
  if (verbose) std::printf("wiener=%g\n", r.wiener);
  // "gammaApo" has a default value, so this always applies: the output
  // apodization is always mode 2 (cudaSirecon.cpp:1038-1045).
  r.apodizeoutput = 2;
  if (verbose) std::printf("gamma = %g\n", r.apoGamma);
  if (auto s = get("saveprefiltered")) {
    set_str(r.fileSeparated, s->values.back());
    r.bSaveSeparated = 1;
  }
  if (auto s = get("savealignedraw")) {
    set_str(r.fileRawAligned, s->values.back());
    r.bSaveAlignedRaw = 1;
  }
  if (auto s = get("saveoverlaps")) {
    set_str(r.fileOverlaps, s->values.back());
    r.bSaveOverlaps = 1;
  }
  if (auto s = get("k0angles")) {
    // comma-separated, empty tokens skipped, each read with strtod
    std::stringstream ss(s->values.back());
    std::string tok;
    r.nk0angles = 0;
    while (std::getline(ss, tok, ',')) {
      if (tok.empty()) continue;
      if (r.nk0angles >= SR_MAX_LIST) throw ArgError("too many values for option 'k0angles'");
      r.k0angles[r.nk0angles++] = (float)std::strtod(tok.c_str(), nullptr);
    }
  }
}

}  // namespace

void set_str(char* dst, const std::string& s) {
  if (s.size() >= SR_PATH_MAX) throw ArgError("path longer than " + std::to_string(SR_PATH_MAX - 1) + " bytes: " + s);
  std::memcpy(dst, s.c_str(), s.size() + 1);
}

void set_default_params(SR_ReconParams& p, SR_ImageParams& img) {
  std::memset(&p, 0, sizeof(p));
  std::memset(&img, 0, sizeof(img));
  p.k0startangle = 1.57193f;
  p.linespacing = 0.177f;
  p.na = 1.36f;
  p.nimm = 1.515f;
  p.ndirs = 3;
  p.nphases = 3;
  p.norders_output = 0;
  p.bTwolens = 0;
  p.bFastSIM = 0;
  p.bBessel = 0;
  p.BesselNA = 0.45f;
  p.otfcutoff = 0.006f;
  p.zoomfact = 2;
  p.z_zoom = 1;
  p.nzPadTo = 0;
  p.explodefact = 1.0f;
  p.bFilteroverlaps = 1;
  p.recalcarrays = 1;
  p.napodize = 10;
  p.nforceamp = 1;
  p.forceamp[0] = 0.f;
  p.bSearchforvector = 1;
  p.bUseTime0k0 = 1;
  p.apodizeoutput = 0;
  p.bSuppress_singularities = 1;
  p.suppression_radius = 10;
  p.bDampenOrder0 = 0;
  p.bFitallphases = 1;
  p.do_rescale = 1;
  p.equalizez = 0;
  p.equalizet = 0;
  p.bNoKz0 = 0;
  p.bRadAvgOTF = 0;
  p.bOneOTFperAngle = 0;
  p.bFixdrift = 0;
  p.drift_filter_fact = 0.0f;
  p.constbkgd = 0.0f;
  p.bBgInExtHdr = 0;
  p.bUsecorr = 0;
  p.electrons_per_bit = 0.6528f;
  p.readoutNoiseVar = 32.42f;
  p.bMakemodel = 0;
  p.bSaveAlignedRaw = 0;
  p.bSaveSeparated = 0;
  p.bSaveOverlaps = 0;
  p.bWriteTitle = 0;
  p.bTIFF = 0;
  img.nwaves = 1;
  img.ntimes = 1;
}

Parsed parse_argv(int argc, char** argv, SR_ReconParams& r, SR_ImageParams& img) {
  Parsed out;
  set_default_params(r, img);
  Store st;
  parse_cli_tokens(argc, argv, st);
  if (st.count("help") || argc <= 1) {
    out.help = true;
    return out;
  }
  if (st.count("version")) {
    out.version = true;
    return out;
  }
  if (!st.count("otf-file")) throw ArgError("the option '--otf-file' is required but missing");
  std::string cfg;
  notify(st, r, img, &cfg);
  out.config_file = cfg;
  if (!cfg.empty()) {
    std::ifstream in(cfg);
    if (!in) {
      std::printf("can not open config file: %s\nproceed without it\n", cfg.c_str());
      out.config_missing = true;
    } else {
      parse_config_stream(in, st);
      notify(st, r, img, nullptr);
    }
  }
  set_params(st, r, true);
  return out;
}

void parse_config(const std::string& file, SR_ReconParams& r, SR_ImageParams& img) {
  set_default_params(r, img);
  Store st;
  std::ifstream in(file);
  if (in) parse_config_stream(in, st);
  notify(st, r, img, nullptr);
  set_params(st, r, false);
}

std::string usage_text() {
  std::string s =
      "Usage:\n cudasirecon [--options] [option-arguments] input_file(or folder) output_file(or file name pattern) "
      "otf_file\n\n List of options:\n";
  for (const Opt& o : kOpts) {
    std::string name = o.shortc ? std::string("  -") + o.shortc + " [ --" + o.name + " ]" : std::string("  --") + o.name;
    if (takes_value(o.kind)) {
      name += (o.kind == Kind::IntImplicit0 || o.kind == Kind::BoolImplicit0) ? " [=arg(=0)]" : " arg";
      if (o.deflt && o.kind != Kind::Switch && *o.deflt) name += std::string(" (=") + o.deflt + ")";
    }
    if (name.size() < 40) name.resize(40, ' ');
    else name += "\n" + std::string(40, ' ');
    s += name + o.help + "\n";
  }
  return s;
}

}  // namespace cli
}  // namespace simr
