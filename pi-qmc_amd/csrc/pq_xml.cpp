// pq_xml.cpp -- the pi-qmc input file (pimc.xml) read into chain descriptors.
//
// Host-only.  A small XML reader (elements, attributes, comments, the
// declaration; text is ignored) builds a tree, and the mapping follows the
// reference's parsers element by element:
//   <SuperCell>, <Species>, <Temperature>   SimInfoParser.cc:30-105
//   <Action> children                       ActionParser.cc:102-456,759-815
//   <Estimators> children                   EstimatorParser.cc:150-215,477-488,643-725
//   <PIMC> body                             PIMCParser.cc:125-490,560-640
// Attribute values are read as the reference's XMLParser reads them (atoi /
// atof of the text, 0 when absent; a bool is true when the text starts with
// t or T), and values with a unit ("25 A", "10000 K", "1 amu") are scaled to
// Hartree atomic units with the reference's conversion table
// (XMLUnitParser.cc:24-112, Units.cc:9-48).
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/piqmc_hip.h"

// the message pq_last_error returns (pq_capi.cpp)
void pq_set_last_error(const char* msg);

namespace {

int xml_fail(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  pq_set_last_error(buf);
  return PQ_ERR_ARG;
}

struct Node {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<Node> kids;
  int line = 0;
  const std::string* get(const char* a) const {
    for (const auto& kv : attrs)
      if (kv.first == a) return &kv.second;
    return nullptr;
  }
  bool has(const char* a) const { return get(a) != nullptr; }
  std::string str(const char* a) const { const std::string* v = get(a); return v ? *v : std::string(); }
  int i(const char* a) const { const std::string* v = get(a); return v ? std::atoi(v->c_str()) : 0; }
  double d(const char* a) const { const std::string* v = get(a); return v ? std::atof(v->c_str()) : 0.0; }
  bool b(const char* a) const { const std::string* v = get(a); return v && !v->empty() && ((*v)[0] == 't' || (*v)[0] == 'T'); }
};

// ------------------------------------------------------------ XML reader

struct Reader {
  const std::string& s;
  size_t p = 0;
  int line = 1;
  explicit Reader(const std::string& text) : s(text) {}
  bool eof() const { return p >= s.size(); }
  void adv(size_t n) {
    for (size_t k = 0; k < n && p < s.size(); ++k, ++p)
      if (s[p] == '\n') ++line;
  }
  void ws() { while (!eof() && std::isspace((unsigned char)s[p])) adv(1); }
  bool starts(const char* t) const { return s.compare(p, std::strlen(t), t) == 0; }
  bool skip_to(const char* t) {
    const size_t q = s.find(t, p);
    if (q == std::string::npos) return false;
    adv(q - p + std::strlen(t));
    return true;
  }
  std::string name() {
    const size_t q0 = p;
    while (!eof() && (std::isalnum((unsigned char)s[p]) || s[p] == '_' || s[p] == '-' || s[p] == ':' || s[p] == '.')) adv(1);
    return s.substr(q0, p - q0);
  }
};

std::string decode(const std::string& v) {
  std::string o;
  for (size_t k = 0; k < v.size(); ++k) {
    if (v[k] != '&') { o += v[k]; continue; }
    const size_t e = v.find(';', k);
    const std::string ent = e == std::string::npos ? "" : v.substr(k + 1, e - k - 1);
    if (ent == "amp") o += '&';
    else if (ent == "lt") o += '<';
    else if (ent == "gt") o += '>';
    else if (ent == "quot") o += '"';
    else if (ent == "apos") o += '\'';
    else { o += v[k]; continue; }
    k = e;
  }
  return o;
}

// markup the reader steps over: the declaration, comments, DOCTYPE, CDATA
int skip_misc(Reader& r) {
  for (;;) {
    r.ws();
    if (r.starts("<?")) { if (!r.skip_to("?>")) return xml_fail("line %d: unterminated <?", r.line); }
    else if (r.starts("<!--")) { if (!r.skip_to("-->")) return xml_fail("line %d: unterminated comment", r.line); }
    else if (r.starts("<![CDATA[")) { if (!r.skip_to("]]>")) return xml_fail("line %d: unterminated CDATA", r.line); }
    else if (r.starts("<!")) { if (!r.skip_to(">")) return xml_fail("line %d: unterminated <!", r.line); }
    else return PQ_OK;
  }
}

int parse_element(Reader& r, Node& n, int depth) {
  if (depth > 64) return xml_fail("line %d: elements nested too deep", r.line);
  if (r.eof() || r.s[r.p] != '<') return xml_fail("line %d: expected an element", r.line);
  r.adv(1);
  n.line = r.line;
  n.name = r.name();
  if (n.name.empty()) return xml_fail("line %d: element without a name", r.line);
  for (;;) {
    r.ws();
    if (r.eof()) return xml_fail("line %d: <%s> not closed", n.line, n.name.c_str());
    if (r.starts("/>")) { r.adv(2); return PQ_OK; }
    if (r.s[r.p] == '>') { r.adv(1); break; }
    std::string an = r.name();
    if (an.empty()) return xml_fail("line %d: bad attribute in <%s>", r.line, n.name.c_str());
    r.ws();
    if (r.eof() || r.s[r.p] != '=') return xml_fail("line %d: attribute %s without a value", r.line, an.c_str());
    r.adv(1);
    r.ws();
    if (r.eof() || (r.s[r.p] != '"' && r.s[r.p] != '\'')) return xml_fail("line %d: attribute %s not quoted", r.line, an.c_str());
    const char qc = r.s[r.p];
    r.adv(1);
    const size_t e = r.s.find(qc, r.p);
    if (e == std::string::npos) return xml_fail("line %d: attribute %s not closed", r.line, an.c_str());
    for (const auto& kv : n.attrs)
      if (kv.first == an) return xml_fail("line %d: attribute %s given twice", r.line, an.c_str());
    n.attrs.emplace_back(an, decode(r.s.substr(r.p, e - r.p)));
    r.adv(e - r.p + 1);
  }
  for (;;) {
    // text between elements is not part of the input's meaning
    while (!r.eof() && r.s[r.p] != '<') r.adv(1);
    if (int rc = skip_misc(r)) return rc;
    if (r.eof()) return xml_fail("line %d: <%s> never closed", n.line, n.name.c_str());
    if (r.starts("</")) {
      r.adv(2);
      const std::string cn = r.name();
      r.ws();
      if (cn != n.name || r.eof() || r.s[r.p] != '>')
        return xml_fail("line %d: </%s> closes <%s> of line %d", r.line, cn.c_str(), n.name.c_str(), n.line);
      r.adv(1);
      return PQ_OK;
    }
    n.kids.emplace_back();
    if (int rc = parse_element(r, n.kids.back(), depth + 1)) return rc;
  }
}

int parse_document(const std::string& text, Node& root) {
  Reader r(text);
  if (int rc = skip_misc(r)) return rc;
  if (int rc = parse_element(r, root, 0)) return rc;
  if (int rc = skip_misc(r)) return rc;
  r.ws();
  if (!r.eof()) return xml_fail("line %d: content after the root element", r.line);
  return PQ_OK;
}

// ------------------------------------------------------------------ units

// Atomic-unit values of one named unit (Units.cc:9-48: the named unit per
// atomic unit, inverted by getLengthScaleIn / getEnergyScaleIn / getMassScaleIn)
double unit_per_au(const std::string& kind, const std::string& u, bool* found) {
  static const std::map<std::string, double> length = {
      {"nm", 0.05291772086}, {"A", 0.5291772086}, {"pm", 52.91772086}, {"um", 0.05291772086e-3},
      {"mm", 0.05291772086e-6}, {"cm", 0.05291772086e-7}, {"m", 0.05291772086e-9}, {"a0", 1.0}};
  static const std::map<std::string, double> energy = {
      {"kV", 0.0272113845}, {"keV", 0.0272113845}, {"eV", 27.2113845}, {"V", 27.2113845},
      {"meV", 27211.3845}, {"ueV", 27211384.5}, {"K", 3.1577465e5}, {"mK", 3.1577465e8},
      {"uK", 3.1577465e11}, {"nK", 3.1577465e14}, {"Ha", 1.0}, {"Ry", 2.0}, {"cm-1", 219474.6314},
      {"kcal/mol", 627.5095}, {"Hz", 6.579683920e15}, {"kHz", 6.579683920e12}, {"MHz", 6.579683920e9},
      {"GHz", 6.579683920e6}, {"THz", 6.579683920e3}, {"PHz", 6.579683920}};
  static const std::map<std::string, double> mass = {
      {"amu", 9.1093826e-31 / 1.66053886e-27}, {"m_p", 9.1093826e-31 / 1.67262171e-27}, {"m_e", 1.0}};
  const std::map<std::string, double>& m = kind == "length" ? length : kind == "energy" ? energy : mass;
  auto it = m.find(u);
  *found = it != m.end();
  return *found ? it->second : 1.0;
}

// XMLUnitParser::parseUnitAndValue: the value, then after the first space a unit
int unit_value(const Node& n, const char* a, const char* kind, double* out) {
  const std::string* v = n.get(a);
  if (!v) { *out = 0; return PQ_OK; }
  const size_t sp = v->find(' ');
  if (sp == std::string::npos) { *out = std::atof(v->c_str()); return PQ_OK; }
  const double value = std::atof(v->substr(0, sp).c_str());
  std::string u = v->substr(sp + 1);
  bool found = false;
  double per = 1;
  if (!std::strcmp(kind, "time")) {
    // "Ha-1"-style inverse energies (XMLUnitParser::getTimeAttribute)
    per = u.size() > 2 ? 1.0 / unit_per_au("energy", u.substr(0, u.size() - 2), &found) : 1.0;
  } else {
    per = unit_per_au(kind, u, &found);
  }
  // the reference warns and takes 1 for an unknown unit; a silent wrong scale
  // is worse than a refusal here
  if (!found) return xml_fail("line %d: unknown %s unit \"%s\" in %s=\"%s\"", n.line, kind, u.c_str(), a, v->c_str());
  *out = value / per;
  return PQ_OK;
}

// --------------------------------------------------------------- mapping

struct Ctx {
  int ndim = 3;
  std::vector<std::string> species;   // names in order
  const Node* root = nullptr;
};

int species_index(const Ctx& x, const Node& n, const char* a, int* out) {
  const std::string name = n.str(a);
  // SimulationInfo::getSpecies: an empty name binds the first species
  if (name.empty()) { *out = 0; return PQ_OK; }
  for (size_t k = 0; k < x.species.size(); ++k)
    if (x.species[k] == name) { *out = (int)k; return PQ_OK; }
  return xml_fail("line %d: <%s %s=\"%s\">: no such species", n.line, n.name.c_str(), a, name.c_str());
}

const Node* first(const Node& n, const char* name) {
  for (const Node& k : n.kids)
    if (k.name == name) return &k;
  return nullptr;
}

int sim_info(Ctx& x, const Node& root, pq_chain_desc& d) {
  d.ndim = x.ndim;
  d.nspecies = 0;
  for (const Node& k : root.kids) {
    if (k.name != "Species") continue;
    if (d.nspecies >= PQ_MAX_SPECIES) return xml_fail("line %d: more than %d species", k.line, PQ_MAX_SPECIES);
    pq_species_desc& s = d.species[d.nspecies++];
    s.count = k.i("count");
    if (int rc = unit_value(k, "mass", "mass", &s.mass)) return rc;
    s.charge = k.d("charge");
    s.is_static = k.b("isStatic") ? 1 : 0;
    if (k.has("AnisotropicMass") || first(k, "AnisotropicMass")) return xml_fail("line %d: AnisotropicMass is not on the device path", k.line);
    x.species.push_back(k.str("name"));
  }
  if (d.nspecies == 0) return xml_fail("the input has no <Species>");
  const Node* sc = first(root, "SuperCell");
  if (!sc) return xml_fail("the input has no <SuperCell>");
  double a = 0;
  if (int rc = unit_value(*sc, "a", "length", &a)) return rc;
  static const char* dn[3] = {"x", "y", "z"};
  for (int k = 0; k < 3; ++k) d.cell[k] = k < x.ndim ? a * sc->d(dn[k]) : 1.0;
  const Node* tn = first(root, "Temperature");
  if (!tn) return xml_fail("the input has no <Temperature>");
  double T = 0;
  if (int rc = unit_value(*tn, "value", "energy", &T)) return rc;
  d.nslice = tn->i("nslice");
  if (d.nslice > 0) {
    d.tau = 1.0 / (T * d.nslice);
  } else {
    if (int rc = unit_value(*tn, "tau", "time", &d.tau)) return rc;
    d.nslice = (int)(1.0 / (T * d.tau) + 0.1);
  }
  return PQ_OK;
}

int actions(Ctx& x, const Node& root, pq_chain_desc& d) {
  const Node* an = first(root, "Action");
  if (!an) return PQ_OK;
  for (const Node& k : an->kids) {
    if (k.name == "SpringAction") {
      d.spring = 1;
    } else if (k.name == "SHOAction") {
      if (d.nsho >= PQ_MAX_SHO) return xml_fail("line %d: more than %d SHOActions", k.line, PQ_MAX_SHO);
      pq_sho_desc& s = d.sho[d.nsho++];
      if (int rc = unit_value(k, "omega", "energy", &s.omega)) return rc;
      if (s.omega == 0) s.omega = 1;
      static const char* dn[3] = {"x", "y", "z"};
      for (int q = 0; q < 3; ++q) {
        s.center[q] = 0;
        if (q < x.ndim)
          if (int rc = unit_value(k, dn[q], "length", &s.center[q])) return rc;
      }
      s.ndim = k.i("ndim");
      if (int rc = species_index(x, k, "species", &s.species)) return rc;
    } else if (k.name == "CoulombAction") {
      if (d.coulomb.enabled) return xml_fail("line %d: a second CoulombAction", k.line);
      pq_coulomb_desc& c = d.coulomb;
      c.enabled = 1;
      c.epsilon = k.d("epsilon") == 0 ? 1.0 : k.d("epsilon");
      c.norder = k.i("norder");
      if (int rc = unit_value(k, "rmin", "length", &c.rmin)) return rc;
      if (int rc = unit_value(k, "rmax", "length", &c.rmax)) return rc;
      c.ngpts = k.i("ngridPoints");
      c.use_ewald = k.b("useEwald") ? 1 : 0;
      std::string et = k.str("ewaldType");
      if (et == "" || et == "opt" || et == "optEwald") c.ewald_type = 0;
      else if (et == "trad" || et == "tradEwald") c.ewald_type = 1;
      else return xml_fail("line %d: ewaldType \"%s\" is not on the device path", k.line, et.c_str());
      if (int rc = unit_value(k, "ewaldRcut", "length", &c.ewald_rcut)) return rc;
      c.ewald_kcut = k.d("ewaldKcut");
      c.ewald_kappa = k.d("kappa");
      const int endim = k.i("ewaldNDim");
      if (c.use_ewald && endim != 0 && endim != x.ndim)
        return xml_fail("line %d: ewaldNDim %d below NDIM (image pair actions) is not on the device path", k.line, endim);
      if (c.use_ewald && k.i("ewaldImages") > 1)
        return xml_fail("line %d: ewaldImages > 1 is not on the device path", k.line);
      if (k.has("exchangeLevel")) return xml_fail("line %d: exchange pair actions (exchangeLevel) are not on the device path", k.line);
      if (k.d("screenDist") != 0) return xml_fail("line %d: screenDist is not on the device path", k.line);
    } else if (k.name == "FixedNodeAction") {
      if (d.nnodes >= PQ_MAX_NODES) return xml_fail("line %d: more than %d FixedNodeActions", k.line, PQ_MAX_NODES);
      pq_node_desc& f = d.nodes[d.nnodes++];
      if (int rc = species_index(x, k, "species", &f.species)) return rc;
      const std::string model = k.str("model");
      if (model == "SHONodes") {
        f.model = PQ_NODES_SHO;
        if (int rc = unit_value(k, "omega", "energy", &f.omega)) return rc;
      } else if (model == "" || model == "FreeParticleNodes") {
        // ActionParser::parseNodeModel: every other name builds FreeParticleNodes;
        // the named models beside SHONodes are not on the device path
        f.model = PQ_NODES_FREE;
      } else if (model == "GSSNode" || model == "GroundStateSNode" || model == "WireNodes" ||
                 model == "ExcitonNodes" || model == "AugmentedNodes") {
        return xml_fail("line %d: node model %s is not on the device path", k.line, model.c_str());
      } else {
        f.model = PQ_NODES_FREE;
      }
      if (int rc = unit_value(k, "temperature", "energy", &f.temperature)) return rc;
      f.nodal_action = k.b("noNodalAction") ? 0 : 1;
      f.many_body_distance = k.b("useManyBodyDistance") ? 1 : 0;
      if (k.b("useDistDerivative")) return xml_fail("line %d: useDistDerivative is not on the device path", k.line);
      if (f.model == PQ_NODES_FREE && (k.b("useHungarian") || k.i("useIterations") > 0))
        return xml_fail("line %d: FreeParticleNodes with useHungarian or useIterations is not on the device path",
                        k.line);
      // FreeParticleNodes::MatrixUpdate (ActionParser.cc:808-811); SHONodes
      // has no useUpdates attribute
      f.use_updates = (f.model == PQ_NODES_FREE && k.b("useUpdates")) ? 1 : 0;
    } else {
      return xml_fail("line %d: <%s> is not on the device path", k.line, k.name.c_str());
    }
  }
  return PQ_OK;
}

int estimators(Ctx& x, const Node& root, pq_xml_input& in) {
  const Node* en = first(root, "Estimators");
  if (!en) return PQ_OK;
  for (const Node& k : en->kids) {
    if (k.name == "ThermalEnergyEstimator") {
      in.thermal_energy = 1;
    } else if (k.name == "VirialEnergyEstimator") {
      in.virial_nwindow = k.i("nwindow") ? k.i("nwindow") : 1;
    } else if (k.name == "CoulombEnergyEstimator") {
      in.coulomb_energy = 1;
      in.coulomb_epsilon = k.d("epsilon") == 0 ? 1.0 : k.d("epsilon");
      in.coulomb_ewald = k.b("useEwald") ? 1 : 0;
      if (in.coulomb_ewald) {
        if (int rc = unit_value(k, "rcut", "length", &in.coulomb_rcut)) return rc;
        in.coulomb_kcut = k.d("kcut");
        const std::string et = k.str("ewaldType");
        in.coulomb_ewald_type = (et == "trad" || et == "tradEwald") ? 1 : 0;
        in.coulomb_kappa = k.d("kappa");
        if (in.coulomb_ewald_type == 1 && k.i("ewaldImages") > 1)
          return xml_fail("line %d: ewaldImages > 1 is not on the device path", k.line);
      }
    } else if (k.name == "PairCFEstimator") {
      if (k.kids.size() != 1 || k.kids[0].name != "Radial" || k.kids[0].has("dir"))
        return xml_fail("line %d: only the one-distance PairCFEstimator with <Radial> is on the device path", k.line);
      int ns = k.i("nspecies");
      if (ns == 0) ns = 2;
      if (ns > PQ_XML_MAX_CF_SPECIES) return xml_fail("line %d: nspecies above %d", k.line, PQ_XML_MAX_CF_SPECIES);
      in.pair_cf_nspecies = ns;
      for (int s = 0; s < ns; ++s) {
        char an[16];
        std::snprintf(an, sizeof an, "species%d", s + 1);
        if (int rc = species_index(x, k, an, &in.pair_cf_species[s])) return rc;
      }
      const Node& r = k.kids[0];
      if (int rc = unit_value(r, "min", "length", &in.pair_cf_rmin)) return rc;
      if (int rc = unit_value(r, "max", "length", &in.pair_cf_rmax)) return rc;
      in.pair_cf_nbin = r.i("nbin");
    } else {
      return xml_fail("line %d: <%s> is not on the device path", k.line, k.name.c_str());
    }
  }
  return PQ_OK;
}

bool has_measure(const Node& n) {
  for (const Node& k : n.kids)
    if (k.name == "Measure") return true;
  return false;
}

bool contains_measure(const Node& n) {
  if (n.name == "Measure") return true;
  for (const Node& k : n.kids)
    if (contains_measure(k)) return true;
  return false;
}

// <Sample> (PIMCParser.cc:312-490) into a sample descriptor and its mover
int sample(const Ctx& x, const Node& k, pq_sample_desc& s, int* mover) {
  const std::string mv = k.str("mover");
  if (mv == "Free") *mover = PQ_MOVER_FREE;
  else if (mv == "FreePBC" || mv == "") *mover = PQ_MOVER_FREE_PBC;
  else return xml_fail("line %d: mover \"%s\" is not on the device path", k.line, mv.c_str());
  const std::string ch = k.str("chooser");
  if (ch == "twoPair" || ch == "pair") return xml_fail("line %d: chooser \"%s\" is not on the device path", k.line, ch.c_str());
  if (k.b("both")) return xml_fail("line %d: both=\"true\" is not on the device path", k.line);
  s.nmoving = k.i("npart");
  const std::string sp = k.str("species");
  const bool noperm = k.b("noPermutation");
  if (sp == "" || sp == "all") {
    s.species = 0;
    s.chooser = noperm ? PQ_CHOOSER_ALL : PQ_CHOOSER_ALL_RANDOM;
  } else {
    if (int rc = species_index(x, k, "species", &s.species)) return rc;
    s.chooser = noperm ? PQ_CHOOSER_SPECIES : PQ_CHOOSER_WALKING;
  }
  s.nrepeat = k.i("nrepeat") == 0 ? 1 : k.i("nrepeat");
  s.delayed_rejection = k.b("delayedRejection") ? 1 : 0;
  s.sigma_factor = k.d("defaultSigmaFactor");
  s.new_sigma_factor = k.d("newSigmaFactor");
  return PQ_OK;
}

struct Group {
  int nlevel, nsample, nsection;
  pq_sample_desc smp[PQ_MAX_SAMPLES];
};

int section(const Ctx& x, const Node& cs, int nsection, Group& g, int* mover) {
  g.nlevel = cs.i("nlevel");
  g.nsection = nsection;
  g.nsample = 0;
  for (const Node& k : cs.kids) {
    if (k.name != "Sample") return xml_fail("line %d: <%s> inside <ChooseSection> is not on the device path", k.line, k.name.c_str());
    if (g.nsample >= PQ_MAX_SAMPLES) return xml_fail("line %d: more than %d Samples in a section", k.line, PQ_MAX_SAMPLES);
    std::memset(&g.smp[g.nsample], 0, sizeof(pq_sample_desc));
    int mv = 0;
    if (int rc = sample(x, k, g.smp[g.nsample], &mv)) return rc;
    if (*mover >= 0 && mv != *mover)
      return xml_fail("line %d: Samples with different movers in one sweep are not on the device path", k.line);
    *mover = mv;
    ++g.nsample;
  }
  if (g.nsample == 0) return xml_fail("line %d: <ChooseSection> without a <Sample>", cs.line);
  return PQ_OK;
}

// a sweep body: <ChooseSection> elements, and <Loop>s of one <ChooseSection>,
// in order; everything else in it ends the sweep at a <Measure> or is refused
int sweep_body(const Ctx& x, const Node& body, std::vector<Group>& gs, int* mover) {
  for (const Node& k : body.kids) {
    if (k.name == "Measure" || k.name == "Collect") continue;
    if (k.name == "ChooseSection") {
      gs.emplace_back();
      if (int rc = section(x, k, 1, gs.back(), mover)) return rc;
    } else if (k.name == "Loop") {
      const Node* cs = nullptr;
      for (const Node& q : k.kids) {
        if (q.name != "ChooseSection" || cs)
          return xml_fail("line %d: a <Loop> inside a sweep must hold one <ChooseSection> only", k.line);
        cs = &q;
      }
      if (!cs) return xml_fail("line %d: empty <Loop> inside a sweep", k.line);
      gs.emplace_back();
      if (int rc = section(x, *cs, k.i("nrepeat"), gs.back(), mover)) return rc;
    } else if (k.name == "WritePaths" || k.name == "Comment") {
      continue;
    } else {
      return xml_fail("line %d: <%s> inside a sweep is not on the device path", k.line, k.name.c_str());
    }
  }
  if (gs.empty()) return xml_fail("line %d: a sweep without a <ChooseSection>", body.line);
  if ((int)gs.size() > PQ_MAX_GROUPS)
    return xml_fail("line %d: more than %d <ChooseSection> groups in a sweep", body.line, PQ_MAX_GROUPS);
  return PQ_OK;
}

void put_groups(const std::vector<Group>& gs, int mover, pq_chain_desc& d) {
  d.mover = mover < 0 ? PQ_MOVER_FREE_PBC : mover;
  d.nlevel = gs[0].nlevel;
  d.nsample = gs[0].nsample;
  std::memcpy(d.sample, gs[0].smp, sizeof d.sample);
  d.nsection = gs[0].nsection;
  d.ngroup = (int)gs.size() - 1;
  for (size_t g = 1; g < gs.size(); ++g) {
    pq_section_group& G = d.group[g - 1];
    G.nlevel = gs[g].nlevel;
    G.nsample = gs[g].nsample;
    std::memcpy(G.sample, gs[g].smp, sizeof G.sample);
    G.nsection = gs[g].nsection;
  }
}

// the measured loop: the first <Loop> whose own body holds a <Measure>; its
// parent <Loop> with a <Collect> gives the number of blocks
const Node* find_measured(const Node& n, const Node** parent) {
  for (const Node& k : n.kids) {
    if (k.name != "Loop") continue;
    if (has_measure(k)) { *parent = &n; return &k; }
    if (const Node* m = find_measured(k, parent)) return m;
  }
  return nullptr;
}

int pimc(Ctx& x, const Node& root, const pq_chain_desc& base, pq_xml_input& in) {
  const Node* pn = first(root, "PIMC");
  if (!pn) return xml_fail("the input has no <PIMC>");
  const Node* parent = nullptr;
  const Node* ml = find_measured(*pn, &parent);
  if (!ml) return xml_fail("line %d: no <Loop> with a <Measure> in <PIMC>", pn->line);
  in.nmeasure = ml->i("nrepeat");
  in.nblock = 1;
  if (parent && parent != pn && parent->name == "Loop" && first(*parent, "Collect")) in.nblock = parent->i("nrepeat");
  std::vector<Group> gs;
  int mover = -1;
  if (int rc = sweep_body(x, *ml, gs, &mover)) return rc;
  in.sweep = base;
  put_groups(gs, mover, in.sweep);
  // top-level elements before the measured loop: the seed, the initial
  // paths and the thermalising loop (its body is the sweep it repeats)
  for (const Node& k : pn->kids) {
    if (contains_measure(k)) break;
    if (k.name == "RandomGenerator") {
      in.has_seed = 1;
      in.iseed = k.i("iseed");
    } else if (k.name == "ReadPaths") {
      const std::string f = k.str("file");
      if (f.size() >= sizeof in.read_paths) return xml_fail("line %d: ReadPaths file name too long", k.line);
      std::snprintf(in.read_paths, sizeof in.read_paths, "%s", f.c_str());
      in.read_bfactor = k.i("bfactor") ? k.i("bfactor") : 1;
    } else if (k.name == "Loop" && !in.has_thermalize) {
      std::vector<Group> tg;
      int tm = -1;
      if (int rc = sweep_body(x, k, tg, &tm)) return rc;
      in.thermalize = base;
      put_groups(tg, tm, in.thermalize);
      in.has_thermalize = 1;
      in.nthermalize = k.i("nrepeat");
    } else if (k.name == "Loop") {
      return xml_fail("line %d: a second thermalising <Loop> is not modelled", k.line);
    } else {
      ++in.nignored;     // path initialisers (SetCubicLattice, ...) and file writers
    }
  }
  return PQ_OK;
}

}  // namespace

extern "C" int pq_parse_xml(const char* text, int ndim, pq_xml_input* out) {
  if (!text || !out) return xml_fail("pq_parse_xml: null argument");
  if (ndim != 2 && ndim != 3) return xml_fail("pq_parse_xml: ndim must be 2 or 3 (got %d)", ndim);
  std::memset(out, 0, sizeof *out);
  Node root;
  if (int rc = parse_document(text, root)) return rc;
  if (root.name != "Simulation") return xml_fail("line %d: root element <%s>, not <Simulation>", root.line, root.name.c_str());
  Ctx x;
  x.ndim = ndim;
  x.root = &root;
  pq_chain_desc base;
  std::memset(&base, 0, sizeof base);
  if (int rc = sim_info(x, root, base)) return rc;
  if (int rc = actions(x, root, base)) return rc;
  if (int rc = estimators(x, root, *out)) return rc;
  if (int rc = pimc(x, root, base, *out)) return rc;
  return PQ_OK;
}

extern "C" int pq_read_xml(const char* filename, int ndim, pq_xml_input* out) {
  if (!filename || !out) return xml_fail("pq_read_xml: null argument");
  std::ifstream f(filename, std::ios::binary);
  if (!f) return xml_fail("cannot open %s", filename);
  std::stringstream ss;
  ss << f.rdbuf();
  return pq_parse_xml(ss.str().c_str(), ndim, out);
}
