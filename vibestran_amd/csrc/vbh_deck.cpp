// vibestran_amd/csrc/vbh_deck.cpp -- native Nastran BDF reader for the cards
// the linear-static path uses (host C++, no device code).
//
// Restates the reference's BdfParser (src/io/bdf_parser.cpp) for:
//   sections         BEGIN BULK / ENDDATA, case control SUBCASE, LOAD =, SPC =,
//                    DISPLACEMENT / STRESS requests (93-395)
//   logical cards    comment and continuation rules (425-533): '+' lines, bare
//                    '*' lines (large field), lines whose first 8 columns are
//                    blank; free field when the line has a comma, large field
//                    when it starts with '*' or its keyword field ends in '*';
//                    continuation fields 1..8 appended, free-field trailing
//                    '+...' markers dropped (709-778)
//   numbers          parse_double: D -> E, an implicit exponent sign gets an
//                    'E' (780-812); parse_int as std::stoi (814-821)
//   cards            GRID (1091-1102), MAT1 (1104-1120), PSHELL (1265-1287),
//                    PSOLID (1289-1306), CQUAD4 (1433-1451), CHEXA / CTETRA /
//                    CPENTA (1473-1532), FORCE / MOMENT (1663-1689), GRAV
//                    (1717-1730), PLOAD4 (1835-1935), SPC / SPC1 (1937-1995),
//                    PARAM SHELLFORM (574-586, 1997-), CORD2R/C/S (2008-2027),
//                    CORD1R/C (2029-2043), MPC (2047-2068), MPCADD (2070-2095),
//                    RBE2 / RBE3 (2100-2183) expanded into MPC equations;
//                    case control MPC = (242-248)
//   coordinates      Model::resolve_coordinates (src/core/model.cpp:175-250):
//                    CORD1x from basic node positions, CORD2x through their
//                    RID chains, GRID CP positions into the basic frame
//                    (build_axes / to_basic, src/core/coord_sys.cpp:10-69); the
//                    node's CD frame as its rotation matrix at the node
//                    (rotation_matrix, coord_sys.cpp:71-146), FORCE / MOMENT
//                    in a CID frame rotated into basic (linear_static.cpp:351-392)
// and builds the flat arrays the device path takes: nodes in ascending ID
// order (the DofMap order), elements in file order with per-element E, nu,
// rho and PSHELL T, and per subcase its SPC masks, nodal FORCE/MOMENT
// components, PLOAD4 faces, GRAV vector and MPC equations.
//
// Large decks (192k MAT1 + PSOLID pairs at 605k DOF) are the reason this is
// native: the text is cut into chunks at card boundaries and the chunks are
// tokenised and converted in parallel, then merged in file order.
//
// Not covered (error, not silently wrong): GRAV in a CID frame, CHEXA20 /
// CTETRA10 / CPENTA15, PLOAD4 direction vectors, mixed PSOLID ISOP within one
// deck, SPC enforced values, a GRID CD naming an undefined frame. Other cards
// are counted as unsupported and skipped, as the reference logs them.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <string_view>
#include <thread>
#include <unordered_map>
#include <vector>

#include <array>
#include <memory>
#include <tuple>

#include "vbh_error.hpp"

namespace vbh {
namespace {

struct DeckError {
  std::string msg;
};
[[noreturn]] void deck_fail(int line, const std::string& m) {
  throw DeckError{"Line " + std::to_string(line) + ": " + m};
}

std::string_view trim(std::string_view s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == ' ' || s[a] == '\t')) ++a;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
  return s.substr(a, b - a);
}

bool is_blank(std::string_view s) {
  for (char c : s)
    if (c != ' ' && c != '\t' && c != '\r') return false;
  return true;
}

std::string upper(std::string_view s) {
  std::string u(s);
  for (auto& c : u) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  return u;
}

// parse_double (bdf_parser.cpp:780-812): blank -> 0; D/d -> E; the first
// '+'/'-' after position 0 not preceded by E/e gets an 'E'; then the longest
// numeric prefix (std::stod semantics).
double parse_double(std::string_view s, int line) {
  if (s.empty()) return 0.0;
  char buf[64];
  if (s.size() > 60) deck_fail(line, "cannot parse double: '" + std::string(s) + "'");
  size_t n = 0;
  for (char c : s) buf[n++] = (c == 'D' || c == 'd') ? 'E' : c;
  buf[n] = 0;
  for (size_t i = 1; i < n; ++i) {
    if ((buf[i] == '+' || buf[i] == '-') &&
        std::toupper(static_cast<unsigned char>(buf[i - 1])) != 'E') {
      std::memmove(buf + i + 1, buf + i, n - i + 1);
      buf[i] = 'E';
      ++n;
      break;
    }
  }
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(buf, &end);
  if (end == buf || errno == ERANGE)
    deck_fail(line, "cannot parse double: '" + std::string(s) + "'");
  return v;
}

// parse_int (bdf_parser.cpp:814-821): blank -> 0, else std::stoi semantics
// (optional sign, digits, the rest ignored; no digits or overflow is an error).
int parse_int(std::string_view s, int line) {
  if (s.empty()) return 0;
  char buf[64];
  if (s.size() > 60) deck_fail(line, "cannot parse int: '" + std::string(s) + "'");
  std::memcpy(buf, s.data(), s.size());
  buf[s.size()] = 0;
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(buf, &end, 10);
  if (end == buf || errno == ERANGE || v < INT32_MIN || v > INT32_MAX)
    deck_fail(line, "cannot parse int: '" + std::string(s) + "'");
  return static_cast<int>(v);
}

bool integer_like(std::string_view s) {  // field_is_integer_like (844-856)
  if (s.empty()) return false;
  size_t i = 0;
  if (s[0] == '+' || s[0] == '-') ++i;
  if (i >= s.size()) return false;
  for (; i < s.size(); ++i)
    if (!std::isdigit(static_cast<unsigned char>(s[i]))) return false;
  return true;
}

int dof_mask(int c) {  // DofSet::from_int (types.hpp:114-124)
  int m = 0;
  for (int t = c; t > 0; t /= 10) {
    const int d = t % 10;
    if (d >= 1 && d <= 6) m |= 1 << (d - 1);
  }
  return m;
}

// ---- records produced by the chunk workers -----------------------------------
struct GridRec { int id, cp, cd, line; double x, y, z; };
struct CoordRec {  // CORD2x: defining points A, B, C in the RID frame; CORD1x: node IDs
  int id, rid, type, line;  // type 0 rectangular, 1 cylindrical, 2 spherical
  bool cord1;
  double a[3], b[3], c[3];
  int na, nb, nc;
};
struct MpcTerm { int node, dof; double coeff; };
struct MpcRec {  // MPC equation, or MPCADD (add = true: sid <- each of srcs)
  bool add;
  int sid, line;
  std::vector<MpcTerm> terms;
  std::vector<int> srcs;
};
struct MatRec { int id; double E, nu, rho; };
struct PropRec { int id, kind, mid; double t; int isop; };  // kind 0 PSOLID, 1 PSHELL
struct ElemRec { int id, type, pid, nn; int nodes[8]; int line; };
struct NodalRec { int sid, node, off, cid; double v[3]; };  // off 0 FORCE, 3 MOMENT
struct GravRec { int sid; double a[3]; };
struct PloadRec { int sid, eid, g1, g34; double p[4]; int line; };
struct SpcRec { int sid, node, mask; };
struct RbeGroup { double w; int comp; std::vector<int> nodes; };
struct RbeRec {  // RBE2: gn, cm, gm; RBE3: ref node, refc, weight groups
  int kind, eid, line, node, comp;
  std::vector<int> gm;
  std::vector<RbeGroup> groups;
};

struct Chunk {
  std::vector<GridRec> grids;
  std::vector<MatRec> mats;
  std::vector<PropRec> props;
  std::vector<ElemRec> elems;
  std::vector<NodalRec> nodal;
  std::vector<GravRec> grav;
  std::vector<PloadRec> pload;
  std::vector<SpcRec> spc;
  std::vector<CoordRec> coords;
  std::vector<MpcRec> mpcs;
  std::vector<RbeRec> rbes;
  std::vector<std::pair<std::string, std::string>> params;
  std::set<std::string> unsupported;
  std::string error;
};

// ---- line classification (bdf_parser.cpp:425-460) ----------------------------
bool line_is_comment(std::string_view l) {
  size_t f = 0;
  while (f < l.size() && (l[f] == ' ' || l[f] == '\t')) ++f;
  return f == l.size() || l[f] == '$' || is_blank(l);
}

bool line_is_continuation(std::string_view l) {
  if (l.empty()) return false;
  const char c0 = l[0];
  if (c0 == '+') return true;
  if (c0 == '*') {
    for (size_t i = 1; i < std::min<size_t>(l.size(), 8); ++i)
      if (!std::isspace(static_cast<unsigned char>(l[i]))) return false;
    return true;
  }
  if (std::isspace(static_cast<unsigned char>(c0))) {
    const size_t kend = std::min<size_t>(l.size(), 8);
    for (size_t i = 0; i < kend; ++i)
      if (!std::isspace(static_cast<unsigned char>(l[i]))) return false;
    for (size_t i = kend; i < l.size(); ++i)
      if (!std::isspace(static_cast<unsigned char>(l[i]))) return true;
  }
  return false;
}

// ---- field splitting (709-778) -------------------------------------------------
void split_small(std::string_view l, std::vector<std::string_view>& f) {
  auto get = [&](size_t start, size_t width) -> std::string_view {
    if (start >= l.size()) return {};
    return trim(l.substr(start, std::min(width, l.size() - start)));
  };
  f.push_back(get(0, 8));
  for (int i = 0; i < 8; ++i) f.push_back(get(8 + 8 * i, 8));
}

void split_large(std::string_view l, std::vector<std::string_view>& f) {
  auto get = [&](size_t start, size_t width) -> std::string_view {
    if (start >= l.size()) return {};
    return trim(l.substr(start, std::min(width, l.size() - start)));
  };
  f.push_back(get(0, 8));
  for (int i = 0; i < 4; ++i) f.push_back(get(8 + 16 * i, 16));
}

void split_free(std::string_view l, std::vector<std::string_view>& f) {
  const size_t d = l.find('$');
  if (d != std::string_view::npos) l = l.substr(0, d);
  const size_t first = f.size();
  // tokens as std::getline(ss, tok, ',') yields them: no token after a final ','
  size_t a = 0;
  while (a < l.size()) {
    const size_t c = l.find(',', a);
    if (c == std::string_view::npos) {
      f.push_back(trim(l.substr(a)));
      break;
    }
    f.push_back(trim(l.substr(a, c - a)));
    a = c + 1;
  }
  // a 10th field starting with '+' or '*' is a continuation label
  if (f.size() - first >= 10) {
    const std::string_view f9 = f[first + 9];
    if (!f9.empty() && (f9[0] == '+' || f9[0] == '*')) f.erase(f.begin() + first + 9);
  }
}

void split_any(std::string_view l, std::vector<std::string_view>& f) {
  if (l.find(',') != std::string_view::npos) {
    split_free(l, f);
    return;
  }
  bool large = !l.empty() && l[0] == '*';
  if (!large) {
    const size_t kend = std::min<size_t>(l.size(), 8);
    size_t last = std::string_view::npos;
    for (size_t i = 0; i < kend; ++i)
      if (l[i] != ' ' && l[i] != '\t') last = i;
    if (last != std::string_view::npos && l[last] == '*') large = true;
  }
  if (large) split_large(l, f);
  else split_small(l, f);
}

std::string_view fld(const std::vector<std::string_view>& f, size_t i) {
  return i < f.size() ? f[i] : std::string_view{};
}

// ---- card processors -------------------------------------------------------------
void process(const std::vector<std::string_view>& f, int line, Chunk& ck) {
  std::string kw = upper(f[0]);
  if (!kw.empty() && kw[0] == '*') kw.erase(0, 1);
  while (!kw.empty() && (kw.back() == '*' || std::isspace(static_cast<unsigned char>(kw.back()))))
    kw.pop_back();
  if (kw == "GRID") {
    const int id = parse_int(fld(f, 1), line);
    const int cp = fld(f, 2).empty() ? 0 : parse_int(fld(f, 2), line);
    const double x = parse_double(fld(f, 3), line), y = parse_double(fld(f, 4), line),
                 z = parse_double(fld(f, 5), line);
    const int cd = fld(f, 6).empty() ? 0 : parse_int(fld(f, 6), line);
    ck.grids.push_back({id, cp, cd, line, x, y, z});
  } else if (kw == "MAT1") {
    ck.mats.push_back({parse_int(fld(f, 1), line), parse_double(fld(f, 2), line),
                       fld(f, 4).empty() ? 0.0 : parse_double(fld(f, 4), line),
                       fld(f, 5).empty() ? 0.0 : parse_double(fld(f, 5), line)});
  } else if (kw == "PSOLID") {
    int isop = -1;
    if (!fld(f, 6).empty()) {
      const std::string u = upper(fld(f, 6));
      isop = (u == "SRI") ? 1 : (u == "EAS" ? 0 : -1);
    }
    ck.props.push_back({parse_int(fld(f, 1), line), 0, parse_int(fld(f, 2), line), 0.0, isop});
  } else if (kw == "PSHELL") {
    ck.props.push_back({parse_int(fld(f, 1), line), 1, parse_int(fld(f, 2), line),
                        parse_double(fld(f, 3), line), -1});
  } else if (kw == "CQUAD4" || kw == "CHEXA" || kw == "CTETRA" || kw == "CPENTA") {
    ElemRec e{};
    e.id = parse_int(fld(f, 1), line);
    e.pid = parse_int(fld(f, 2), line);
    e.line = line;
    if (kw == "CQUAD4") {
      e.type = VBH_CQUAD4;
      e.nn = 4;
      for (int i = 0; i < 4; ++i) e.nodes[i] = parse_int(fld(f, 3 + i), line);
    } else {
      int nn = 0;
      for (size_t i = 3; i < f.size(); ++i) {
        if (f[i].empty()) continue;
        if (nn == 8) {
          nn = 9;
          break;
        }
        e.nodes[nn++] = parse_int(f[i], line);
      }
      const int want = kw == "CHEXA" ? 8 : kw == "CPENTA" ? 6 : 4;
      if (nn != want)
        deck_fail(line, kw + " has " + (nn > 8 ? std::string("more than 8") : std::to_string(nn)) +
                            " nodes; expected " + std::to_string(want));
      e.nn = nn;
      e.type = kw == "CHEXA" ? VBH_CHEXA8 : kw == "CPENTA" ? VBH_CPENTA6 : VBH_CTETRA4;
    }
    ck.elems.push_back(e);
  } else if (kw == "FORCE" || kw == "MOMENT") {
    const int sid = parse_int(fld(f, 1), line), g = parse_int(fld(f, 2), line);
    const int cid = fld(f, 3).empty() ? 0 : parse_int(fld(f, 3), line);
    const double s = parse_double(fld(f, 4), line);
    ck.nodal.push_back({sid, g, kw == "FORCE" ? 0 : 3, cid,
                        {s * parse_double(fld(f, 5), line), s * parse_double(fld(f, 6), line),
                         s * parse_double(fld(f, 7), line)}});
  } else if (kw == "GRAV") {
    if (parse_int(fld(f, 2), line) != 0)
      deck_fail(line, "GRAV in a local coordinate system (CID) is not supported");
    for (size_t i = 7; i < f.size(); ++i)
      if (!f[i].empty())
        deck_fail(line, "GRAV rotational acceleration components are not yet supported");
    const double s = fld(f, 3).empty() ? 0.0 : parse_double(fld(f, 3), line);
    ck.grav.push_back({parse_int(fld(f, 1), line),
                       {s * parse_double(fld(f, 4), line), s * parse_double(fld(f, 5), line),
                        s * parse_double(fld(f, 6), line)}});
  } else if (kw == "PLOAD4") {
    const int sid = parse_int(fld(f, 1), line);
    const bool alt = f.size() > 3 && !fld(f, 2).empty() && !integer_like(fld(f, 2)) &&
                     integer_like(fld(f, 3));
    for (size_t i = 9; i < std::min<size_t>(f.size(), 13); ++i)
      if (!f[i].empty()) deck_fail(line, "PLOAD4 direction vectors are not supported");
    double p[4];
    std::vector<int> eids;
    int g1 = -1, g34 = -1;
    const bool thru = upper(fld(f, 7)) == "THRU";
    if (alt) {
      p[0] = p[1] = p[2] = p[3] = parse_double(fld(f, 2), line);
      if (thru) {
        // expand_id_list from field 3: ids with THRU ranges
        int prev = 0;
        bool in_thru = false;
        for (size_t i = 3; i < f.size(); ++i) {
          if (f[i].empty()) continue;
          if (upper(f[i]) == "THRU") {
            in_thru = true;
            continue;
          }
          const int v = parse_int(f[i], line);
          if (in_thru) {
            for (int e = prev + 1; e <= v; ++e) eids.push_back(e);
            in_thru = false;
          } else {
            eids.push_back(v);
          }
          prev = v;
        }
      } else {
        eids.push_back(parse_int(fld(f, 3), line));
      }
    } else {
      if (!integer_like(fld(f, 2)))
        deck_fail(line, "PLOAD4 element field must be an integer element ID");
      p[0] = parse_double(fld(f, 3), line);
      for (int k = 1; k < 4; ++k)
        p[k] = fld(f, 3 + k).empty() ? p[0] : parse_double(fld(f, 3 + k), line);
      if (thru) {
        const int e1 = parse_int(fld(f, 2), line), e2 = parse_int(fld(f, 8), line);
        if (e2 < e1)
          deck_fail(line, "PLOAD4 THRU range end " + std::to_string(e2) +
                              " is less than start " + std::to_string(e1));
        for (int e = e1; e <= e2; ++e) eids.push_back(e);
      } else {
        eids.push_back(parse_int(fld(f, 2), line));
        const bool h7 = !fld(f, 7).empty(), h8 = !fld(f, 8).empty();
        if (h7 && h8) {
          g1 = parse_int(fld(f, 7), line);
          g34 = parse_int(fld(f, 8), line);
        } else if (h7 || h8) {
          deck_fail(line, "PLOAD4 solid face selection requires both face grid fields");
        }
      }
    }
    if (eids.empty()) deck_fail(line, "PLOAD4 requires at least one element ID");
    for (int e : eids) ck.pload.push_back({sid, e, g1, g34, {p[0], p[1], p[2], p[3]}, line});
  } else if (kw == "SPC") {
    const int sid = parse_int(fld(f, 1), line);
    for (int i = 0; i < 2; ++i) {
      const size_t ni = 2 + 3 * i;
      if (ni >= f.size() || f[ni].empty()) break;
      if (!fld(f, ni + 2).empty() && parse_double(fld(f, ni + 2), line) != 0.0)
        deck_fail(line, "SPC enforced displacements are not supported");
      ck.spc.push_back({sid, parse_int(f[ni], line), dof_mask(parse_int(fld(f, ni + 1), line))});
    }
  } else if (kw == "SPC1") {
    const int sid = parse_int(fld(f, 1), line);
    const int mask = dof_mask(parse_int(fld(f, 2), line));
    std::vector<int> ids;
    bool thru = false;
    int start = 0;
    for (size_t i = 3; i < f.size(); ++i) {
      if (f[i].empty()) continue;
      if (upper(f[i]) == "THRU") {
        thru = true;
        if (!ids.empty()) ids.pop_back();
      } else if (thru) {
        const int end = parse_int(f[i], line);
        for (int g = start; g <= end; ++g) ids.push_back(g);
        thru = false;
      } else {
        start = parse_int(f[i], line);
        ids.push_back(start);
      }
    }
    for (int g : ids) ck.spc.push_back({sid, g, mask});
  } else if (kw == "CORD2R" || kw == "CORD2C" || kw == "CORD2S") {
    CoordRec c{};
    c.id = parse_int(fld(f, 1), line);
    c.rid = fld(f, 2).empty() ? 0 : parse_int(fld(f, 2), line);
    c.type = kw[5] == 'R' ? 0 : (kw[5] == 'C' ? 1 : 2);
    c.line = line;
    for (int k = 0; k < 3; ++k) {
      c.a[k] = parse_double(fld(f, 3 + k), line);
      c.b[k] = parse_double(fld(f, 6 + k), line);
      c.c[k] = parse_double(fld(f, 9 + k), line);
    }
    ck.coords.push_back(c);
  } else if (kw == "CORD1R" || kw == "CORD1C") {
    CoordRec c{};
    c.id = parse_int(fld(f, 1), line);
    c.type = kw[5] == 'R' ? 0 : 1;
    c.cord1 = true;
    c.line = line;
    c.na = parse_int(fld(f, 2), line);
    c.nb = parse_int(fld(f, 3), line);
    c.nc = parse_int(fld(f, 4), line);
    ck.coords.push_back(c);
  } else if (kw == "MPC") {
    MpcRec m{};
    m.sid = parse_int(fld(f, 1), line);
    m.line = line;
    for (size_t i = 2; i + 2 < f.size(); i += 3) {
      if (f[i].empty()) break;
      const int g = parse_int(f[i], line), c = parse_int(f[i + 1], line);
      if (c < 1 || c > 6) deck_fail(line, "MPC component " + std::to_string(c) + " is not 1..6");
      m.terms.push_back({g, c, parse_double(f[i + 2], line)});
    }
    if (!m.terms.empty()) ck.mpcs.push_back(std::move(m));
  } else if (kw == "RBE2") {
    // a field with '.', 'E' or 'e' is ALPHA and ends the dependent grid list
    RbeRec r{};
    r.kind = 2;
    r.line = line;
    r.eid = parse_int(fld(f, 1), line);
    r.node = parse_int(fld(f, 2), line);
    r.comp = dof_mask(parse_int(fld(f, 3), line));
    for (size_t i = 4; i < f.size(); ++i) {
      if (f[i].empty()) continue;
      if (f[i].find_first_of(".Ee") != std::string_view::npos) break;
      r.gm.push_back(parse_int(f[i], line));
    }
    ck.rbes.push_back(std::move(r));
  } else if (kw == "RBE3") {
    // groups: a weight, a component set, then grids up to the next real field
    RbeRec r{};
    r.kind = 3;
    r.line = line;
    r.eid = parse_int(fld(f, 1), line);
    r.node = parse_int(fld(f, 3), line);
    r.comp = dof_mask(parse_int(fld(f, 4), line));
    size_t i = 5;
    while (i < f.size()) {
      if (f[i].empty()) {
        ++i;
        continue;
      }
      double w = 0.0;
      try {
        w = parse_double(f[i], line);
      } catch (const DeckError&) {
        ++i;
        continue;
      }
      ++i;
      if (i >= f.size() || f[i].empty()) break;
      RbeGroup g{w, dof_mask(parse_int(f[i], line)), {}};
      ++i;
      while (i < f.size()) {
        if (f[i].empty()) {
          ++i;
          continue;
        }
        if (f[i].find_first_of(".Ee") != std::string_view::npos) break;
        try {
          g.nodes.push_back(parse_int(f[i], line));
        } catch (const DeckError&) {
          break;
        }
        ++i;
      }
      if (!g.nodes.empty()) r.groups.push_back(std::move(g));
    }
    ck.rbes.push_back(std::move(r));
  } else if (kw == "MPCADD") {
    MpcRec m{};
    m.add = true;
    m.sid = parse_int(fld(f, 1), line);
    m.line = line;
    for (size_t i = 2; i < f.size(); ++i)
      if (!f[i].empty()) m.srcs.push_back(parse_int(f[i], line));
    ck.mpcs.push_back(std::move(m));
  } else if (kw == "PARAM") {
    if (f.size() >= 3 && !f[1].empty())
      ck.params.emplace_back(upper(f[1]), upper(f[2]));
  } else if (!kw.empty()) {
    ck.unsupported.insert(kw);
  }
}

// Cards of lines [a, b) (a starts a card or a comment).
void parse_chunk(const std::vector<std::string_view>& lines, size_t a, size_t b, size_t line0,
                 Chunk& ck) {
  std::vector<std::string_view> f;
  f.reserve(64);
  try {
    for (size_t i = a; i < b; ++i) {
      const std::string_view l = lines[i];
      if (line_is_comment(l) || line_is_continuation(l)) continue;
      const size_t first = i;
      f.clear();
      const bool free = l.find(',') != std::string_view::npos;
      split_any(l, f);
      if (free && f.size() > 1 && !f.back().empty() && f.back()[0] == '+') f.pop_back();
      size_t j = i + 1;
      std::vector<std::string_view> cf;
      while (j < b) {
        const std::string_view nx = lines[j];
        if (line_is_comment(nx)) {
          ++j;
          continue;
        }
        if (!line_is_continuation(nx)) break;
        cf.clear();
        split_any(nx, cf);
        const bool nfree = nx.find(',') != std::string_view::npos;
        if (nfree && !cf.empty() && !cf.back().empty() && cf.back()[0] == '+') cf.pop_back();
        for (size_t k = 1; k < cf.size() && k <= 8; ++k) f.push_back(cf[k]);
        ++j;
      }
      i = j - 1;
      while (f.size() < 10) f.push_back({});
      process(f, static_cast<int>(line0 + first + 1), ck);
    }
  } catch (const DeckError& e) {
    ck.error = e.msg;
  }
}

}  // namespace

// ---- the parsed deck -------------------------------------------------------------
struct Deck {
  std::vector<int> node_ids;
  std::vector<double> xyz;
  std::vector<int> elem_ids, etype, eptr, conn;
  std::vector<double> E, nu, rho, thick;
  int solid_form = 0, shell_form = 0;
  std::vector<int> cd;       // CD frame ID per node (0 = basic)
  std::vector<double> cd_T;  // per node: rotation of the CD frame at the node (3x3 row-major)
  struct Sub {
    int id = 1, load = 0, spc = 0, disp = 0, stress = 0, mpc = 0;
    std::vector<int> spc_nodes, spc_masks;
    std::vector<int> mpc_ptr{0}, mpc_node, mpc_dof;  // MPC equations of the MPC set
    std::vector<double> mpc_coeff;
    std::vector<double> nodal;  // nnodes x 6
    std::vector<int> pl_elem, pl_g1, pl_g34;
    std::vector<double> pl_p;
    double grav[3] = {0, 0, 0};
  };
  std::vector<Sub> subs;
  std::vector<std::string> unsupported;
  // RBE2 / RBE3 as MPC equations (every subcase), and the DOFs they activate
  std::vector<int> rigid_ptr{0}, rigid_node, rigid_dof, rigid_active;
  std::vector<double> rigid_coeff;
};

namespace {

// ID -> index lookup: a dense table when the IDs are compact (the usual
// numbering), a hash map otherwise. A later definition of an ID replaces the
// earlier one, as the reference's map assignments do.
class IdIndex {
 public:
  template <class GetId>
  void build(size_t n, GetId id_of) {
    int lo = 0, hi = -1;
    for (size_t i = 0; i < n; ++i) {
      const int id = id_of(i);
      if (i == 0 || id < lo) lo = id;
      if (i == 0 || id > hi) hi = id;
    }
    dense_ = n > 0 && lo >= 0 && static_cast<long long>(hi) < 4ll * static_cast<long long>(n) + 1024;
    if (dense_) {
      table_.assign(static_cast<size_t>(hi) + 1, -1);
      for (size_t i = 0; i < n; ++i) table_[static_cast<size_t>(id_of(i))] = static_cast<int>(i);
    } else {
      map_.reserve(2 * n);
      for (size_t i = 0; i < n; ++i) map_[id_of(i)] = static_cast<int>(i);
    }
  }
  int find(int id) const {  // -1 when absent
    if (dense_) return (id >= 0 && static_cast<size_t>(id) < table_.size()) ? table_[static_cast<size_t>(id)] : -1;
    const auto it = map_.find(id);
    return it == map_.end() ? -1 : it->second;
  }

 private:
  bool dense_ = false;
  std::vector<int> table_;
  std::unordered_map<int, int> map_;
};

// ---- coordinate frames (src/core/coord_sys.cpp) -------------------------------------
struct Frame {
  int type = 0;           // 0 rectangular, 1 cylindrical, 2 spherical
  double o[3] = {0, 0, 0};
  double ax[3][3] = {};   // ax[k] = local axis k in basic
};

void unit(const double* v, double* out, int line) {  // Vec3::normalized (types.hpp:173-178)
  const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (n < 1e-15) deck_fail(line, "Cannot normalize zero vector");
  for (int k = 0; k < 3; ++k) out[k] = v[k] / n;
}

void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// build_axes (coord_sys.cpp:10-30): z along B - A, y = z x (C - A), x = y x z.
void build_axes(Frame& fr, const double* a, const double* b, const double* c, int line) {
  double d[3], ez[3], ac[3], ey[3], ex[3];
  for (int k = 0; k < 3; ++k) d[k] = b[k] - a[k];
  unit(d, ez, line);
  for (int k = 0; k < 3; ++k) d[k] = c[k] - a[k];
  unit(d, ac, line);
  cross(ez, ac, d);
  const double ly = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (ly < 1e-15) deck_fail(line, "Coordinate system defining points are collinear");
  for (int k = 0; k < 3; ++k) ey[k] = d[k] / ly;
  cross(ey, ez, d);
  const double lx = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int k = 0; k < 3; ++k) ex[k] = d[k] / lx;
  for (int k = 0; k < 3; ++k) {
    fr.o[k] = a[k];
    fr.ax[0][k] = ex[k];
    fr.ax[1][k] = ey[k];
    fr.ax[2][k] = ez[k];
  }
}

// to_basic (coord_sys.cpp:32-69): (r, theta deg, z) or (rho, theta deg, phi deg).
void to_basic(const Frame& fr, const double* p, double* out) {
  double q[3] = {p[0], p[1], p[2]};
  constexpr double kDeg = 3.14159265358979323846 / 180.0;
  if (fr.type == 1) {
    const double th = p[1] * kDeg;
    q[0] = p[0] * std::cos(th);
    q[1] = p[0] * std::sin(th);
    q[2] = p[2];
  } else if (fr.type == 2) {
    const double th = p[1] * kDeg, ph = p[2] * kDeg;
    q[0] = p[0] * std::sin(ph) * std::cos(th);
    q[1] = p[0] * std::sin(ph) * std::sin(th);
    q[2] = p[0] * std::cos(ph);
  }
  for (int i = 0; i < 3; ++i)
    out[i] = fr.o[i] + fr.ax[0][i] * q[0] + fr.ax[1][i] * q[1] + fr.ax[2][i] * q[2];
}

// rotation_matrix (coord_sys.cpp:71-146): columns are the frame's local unit
// vectors at basic position pos (for curvilinear frames e_r, e_theta, e_z or
// e_rho, e_phi, e_theta).
void rotation(const Frame& fr, const double* pos, double T[3][3]) {
  double R[3][3];  // R(i, k) = ax[k][i]
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) R[i][k] = fr.ax[k][i];
  if (fr.type == 0) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) T[i][j] = R[i][j];
    return;
  }
  double dp[3], l[3];
  for (int k = 0; k < 3; ++k) dp[k] = pos[k] - fr.o[k];
  for (int k = 0; k < 3; ++k) l[k] = fr.ax[k][0] * dp[0] + fr.ax[k][1] * dp[1] + fr.ax[k][2] * dp[2];
  const double th = std::atan2(l[1], l[0]);
  const double ct = std::cos(th), st = std::sin(th);
  if (fr.type == 1) {
    for (int i = 0; i < 3; ++i) {
      T[i][0] = R[i][0] * ct + R[i][1] * st;
      T[i][1] = R[i][0] * (-st) + R[i][1] * ct;
      T[i][2] = R[i][2];
    }
    return;
  }
  const double ph = std::atan2(std::sqrt(l[0] * l[0] + l[1] * l[1]), l[2]);
  const double cp = std::cos(ph), sp = std::sin(ph);
  for (int i = 0; i < 3; ++i) {
    T[i][0] = R[i][0] * (sp * ct) + R[i][1] * (sp * st) + R[i][2] * cp;
    T[i][1] = R[i][0] * (cp * ct) + R[i][1] * (cp * st) + R[i][2] * (-sp);
    T[i][2] = R[i][0] * (-st) + R[i][1] * ct;
  }
}

// Model::resolve_coordinates (model.cpp:175-250) over the merged records: the
// frames by ID, and GRID CP positions moved into the basic frame in place.
std::map<int, Frame> resolve_frames(const std::vector<CoordRec>& recs, std::vector<GridRec>& grids) {
  std::map<int, const CoordRec*> last;  // a later definition replaces an earlier one
  for (const auto& r : recs) last[r.id] = &r;
  std::map<int, Frame> fr;
  auto raw_node = [&](int id, const CoordRec& r, const char* role) {
    const auto it = std::lower_bound(grids.begin(), grids.end(), id,
                                     [](const GridRec& g, int v) { return g.id < v; });
    if (it == grids.end() || it->id != id)
      deck_fail(r.line, "CORD1x " + std::to_string(r.id) + ": " + role + " node " +
                            std::to_string(id) + " not found");
    return std::array<double, 3>{it->x, it->y, it->z};  // as given: basic
  };
  for (const auto& [id, r] : last) {
    Frame f;
    f.type = r->type;
    if (r->cord1) {
      const auto a = raw_node(r->na, *r, "A"), b = raw_node(r->nb, *r, "B"), c = raw_node(r->nc, *r, "C");
      build_axes(f, a.data(), b.data(), c.data(), r->line);
      fr[id] = f;
    } else if (r->rid == 0) {
      build_axes(f, r->a, r->b, r->c, r->line);
      fr[id] = f;
    }
  }
  for (bool progress = true; progress;) {
    progress = false;
    for (const auto& [id, r] : last) {
      if (fr.count(id)) continue;
      const auto p = fr.find(r->rid);
      if (p == fr.end()) continue;
      double a[3], b[3], c[3];
      to_basic(p->second, r->a, a);
      to_basic(p->second, r->b, b);
      to_basic(p->second, r->c, c);
      Frame f;
      f.type = r->type;
      build_axes(f, a, b, c, r->line);
      fr[id] = f;
      progress = true;
    }
  }
  for (const auto& [id, r] : last)
    if (!fr.count(id))
      deck_fail(r->line, "Coordinate system " + std::to_string(id) +
                             ": circular or unresolvable reference chain (RID=" +
                             std::to_string(r->rid) + " is part of a cycle)");
  for (auto& g : grids) {
    if (g.cp == 0) continue;
    const auto it = fr.find(g.cp);
    if (it == fr.end())
      deck_fail(g.line, "Node " + std::to_string(g.id) + ": CP=" + std::to_string(g.cp) +
                            " references undefined coordinate system");
    const double p[3] = {g.x, g.y, g.z};
    double q[3];
    to_basic(it->second, p, q);
    g.x = q[0];
    g.y = q[1];
    g.z = q[2];
    g.cp = 0;
  }
  return fr;
}

void parse_text(std::string_view text, Deck& d) {
  // lines
  std::vector<std::string_view> lines;
  lines.reserve(text.size() / 40 + 16);
  size_t a = 0;
  while (a < text.size()) {
    size_t e = text.find('\n', a);
    if (e == std::string_view::npos) e = text.size();
    std::string_view l = text.substr(a, e - a);
    if (!l.empty() && l.back() == '\r') l.remove_suffix(1);
    lines.push_back(l);
    a = e + 1;
  }
  // sections
  size_t begin_bulk = static_cast<size_t>(-1), enddata = lines.size();
  for (size_t i = 0; i < lines.size(); ++i) {
    std::string_view l = lines[i];
    size_t f = 0;
    while (f < l.size() && (l[f] == ' ' || l[f] == '\t')) ++f;
    if (f == l.size()) continue;
    const char c0 = static_cast<char>(std::toupper(static_cast<unsigned char>(l[f])));
    if (c0 != 'B' && c0 != 'E') continue;  // only BEGIN BULK / ENDDATA matter here
    const std::string kwl = upper(l.substr(f, std::min<size_t>(l.size() - f, 12)));
    if (kwl.rfind("BEGIN BULK", 0) == 0 || kwl.rfind("BEGIN  BULK", 0) == 0) {
      begin_bulk = i + 1;
    } else if (kwl.rfind("ENDDATA", 0) == 0) {
      enddata = i;
      break;
    }
  }
  if (begin_bulk == static_cast<size_t>(-1)) begin_bulk = 0;
  // case control (93-395, the requests the linear-static path reads)
  {
    Deck::Sub global, cur;
    bool in_sc = false;
    for (size_t i = 0; i < begin_bulk; ++i) {
      std::string u = upper(lines[i]);
      const size_t dl = u.find('$');
      if (dl != std::string::npos) u.resize(dl);
      const std::string_view kw = trim(u);
      if (kw.empty()) continue;
      Deck::Sub& sc = in_sc ? cur : global;
      auto value_after_eq = [&](int line) -> int {
        const size_t eq = kw.find('=');
        if (eq == std::string_view::npos) return -1;
        std::string_view v = trim(kw.substr(eq + 1));
        if (v.empty() || !(std::isdigit(static_cast<unsigned char>(v[0])) || v[0] == '+' || v[0] == '-'))
          return -1;
        return parse_int(v, line);
      };
      // DISPLACEMENT / STRESS [(modifiers)] = ALL | NONE | set: NONE clears the
      // print request; with a modifier list, PRINT asks for it; with neither
      // PRINT nor PLOT listed, STRESS still prints (parse_print_plot_modifiers
      // default) and DISPLACEMENT does not; without a list both print.
      auto request = [&](bool& flag_print, bool default_print) {
        const size_t eq = kw.find('=');
        if (eq == std::string_view::npos) return;
        const std::string_view v = trim(kw.substr(eq + 1));
        if (v.rfind("NONE", 0) == 0) {
          flag_print = false;
          return;
        }
        const size_t lp = kw.find('('), rp = kw.find(')');
        if (lp != std::string_view::npos && rp != std::string_view::npos && rp > lp) {
          const std::string_view mods = kw.substr(lp + 1, rp - lp - 1);
          const bool pr = mods.find("PRINT") != std::string_view::npos;
          const bool pl = mods.find("PLOT") != std::string_view::npos;
          if (pr || (default_print && !pl)) flag_print = true;
        } else {
          flag_print = true;
        }
      };
      const int line = static_cast<int>(i + 1);
      try {
        if (kw.rfind("SUBCASE", 0) == 0) {
          if (in_sc) d.subs.push_back(cur);
          cur = global;
          cur.id = 1;
          in_sc = true;
          const size_t eq = kw.find('=');
          std::string_view v = eq != std::string_view::npos ? kw.substr(eq + 1) : kw.substr(7);
          v = trim(v);
          if (!v.empty()) cur.id = parse_int(v, line);
        } else if (kw.rfind("LOAD", 0) == 0) {
          const int v = value_after_eq(line);
          if (v >= 0) sc.load = v;
        } else if (kw.rfind("MPC", 0) == 0 && kw.find('=') != std::string_view::npos &&
                   kw.rfind("MPCADD", 0) != 0) {
          const int v = value_after_eq(line);
          if (v >= 0) sc.mpc = v;
        } else if (kw.rfind("SPC", 0) == 0 && kw.rfind("SPCADD", 0) != 0) {
          const int v = value_after_eq(line);
          if (v >= 0) sc.spc = v;
        } else if (kw.rfind("DISPLACEMENT", 0) == 0) {
          bool f = sc.disp != 0;
          request(f, false);
          sc.disp = f;
        } else if (kw.rfind("STRESS", 0) == 0 || kw.rfind("STRAIN", 0) == 0) {
          bool f = sc.stress != 0;
          request(f, true);
          sc.stress = f;
        }
      } catch (const DeckError&) {
        // malformed numbers in case control are ignored, as the reference's
        // try { std::stoi } catch (...) {} does
      }
    }
    if (in_sc) d.subs.push_back(cur);
    if (d.subs.empty()) {
      const bool any = global.load || global.spc || global.disp || global.stress;
      if (!any) {
        global.load = 1;
        global.spc = 1;
      }
      d.subs.push_back(global);
    }
  }
  // bulk: chunks cut at card starts
  const size_t nb = enddata > begin_bulk ? enddata - begin_bulk : 0;
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t nchunk = std::max<size_t>(1, std::min<size_t>(hw, nb / 20000 + 1));
  std::vector<size_t> cut(nchunk + 1, enddata);
  cut[0] = begin_bulk;
  for (size_t c = 1; c < nchunk; ++c) {
    size_t k = begin_bulk + nb * c / nchunk;
    while (k < enddata && (line_is_comment(lines[k]) || line_is_continuation(lines[k]))) ++k;
    cut[c] = std::max(k, cut[c - 1]);
  }
  std::vector<Chunk> chunks(nchunk);
  {
    std::vector<std::thread> pool;
    for (size_t c = 1; c < nchunk; ++c)
      pool.emplace_back(parse_chunk, std::cref(lines), cut[c], cut[c + 1], 0, std::ref(chunks[c]));
    parse_chunk(lines, cut[0], cut[1], 0, chunks[0]);
    for (auto& t : pool) t.join();
  }
  for (const auto& ck : chunks)
    if (!ck.error.empty()) throw DeckError{ck.error};

  // merge: nodes sorted by ID
  std::vector<GridRec> grids;
  std::vector<MatRec> mats;
  std::vector<PropRec> props;
  std::set<std::string> unsupported;
  std::map<std::string, std::string> params;
  size_t ne = 0;
  for (auto& ck : chunks) {
    grids.insert(grids.end(), ck.grids.begin(), ck.grids.end());
    mats.insert(mats.end(), ck.mats.begin(), ck.mats.end());
    props.insert(props.end(), ck.props.begin(), ck.props.end());
    for (const auto& [k, v] : ck.params) params[k] = v;
    unsupported.insert(ck.unsupported.begin(), ck.unsupported.end());
    ne += ck.elems.size();
  }
  IdIndex mat_ix, prop_ix;
  mat_ix.build(mats.size(), [&](size_t i) { return mats[i].id; });
  prop_ix.build(props.size(), [&](size_t i) { return props[i].id; });
  // a later GRID with the same ID replaces the earlier one (model.nodes[id] = g)
  std::stable_sort(grids.begin(), grids.end(), [](const GridRec& x, const GridRec& y) { return x.id < y.id; });
  std::vector<GridRec> uniq;
  uniq.reserve(grids.size());
  for (const auto& g : grids) {
    if (!uniq.empty() && uniq.back().id == g.id) uniq.back() = g;
    else uniq.push_back(g);
  }
  std::vector<CoordRec> coords;
  for (const auto& ck : chunks) coords.insert(coords.end(), ck.coords.begin(), ck.coords.end());
  const std::map<int, Frame> frames = resolve_frames(coords, uniq);
  const int nn = static_cast<int>(uniq.size());
  d.cd.assign(static_cast<size_t>(nn), 0);
  d.cd_T.assign(9 * static_cast<size_t>(nn), 0.0);
  for (int i = 0; i < nn; ++i) {
    double* T = &d.cd_T[9 * static_cast<size_t>(i)];
    T[0] = T[4] = T[8] = 1.0;
    const GridRec& g = uniq[static_cast<size_t>(i)];
    if (g.cd == 0) continue;
    const auto it = frames.find(g.cd);
    if (it == frames.end())
      deck_fail(g.line, "Node " + std::to_string(g.id) + ": CD=" + std::to_string(g.cd) +
                            " references undefined coordinate system");
    const double pos[3] = {g.x, g.y, g.z};
    double R[3][3];
    rotation(it->second, pos, R);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) T[3 * a + b] = R[a][b];
    d.cd[static_cast<size_t>(i)] = g.cd;
  }
  d.node_ids.resize(nn);
  d.xyz.resize(3 * static_cast<size_t>(nn));
  for (int i = 0; i < nn; ++i) {
    d.node_ids[i] = uniq[i].id;
    d.xyz[3 * i] = uniq[i].x;
    d.xyz[3 * i + 1] = uniq[i].y;
    d.xyz[3 * i + 2] = uniq[i].z;
  }
  IdIndex node_ix;
  node_ix.build(d.node_ids.size(), [&](size_t i) { return d.node_ids[i]; });
  auto node_index = [&](int id, int line) {
    const int k = node_ix.find(id);
    if (k < 0) deck_fail(line, "GRID " + std::to_string(id) + " is not defined");
    return k;
  };
  auto it_shell = params.find("SHELLFORM");
  d.shell_form = (it_shell != params.end() && it_shell->second == "MINDLIN") ? 1 : 0;
  // elements in file order
  d.elem_ids.reserve(ne);
  d.etype.reserve(ne);
  d.eptr.reserve(ne + 1);
  d.eptr.push_back(0);
  d.E.reserve(ne);
  d.nu.reserve(ne);
  d.rho.reserve(ne);
  d.thick.reserve(ne);
  int isop_seen = -2;
  for (const auto& ck : chunks)
    for (const auto& e : ck.elems) {
      const int ip = prop_ix.find(e.pid);
      if (ip < 0) deck_fail(e.line, "property " + std::to_string(e.pid) + " is not defined");
      const PropRec& p = props[static_cast<size_t>(ip)];
      if ((e.type == VBH_CQUAD4) != (p.kind == 1))
        deck_fail(e.line, "element " + std::to_string(e.id) + " references a property of the wrong kind");
      const int im = mat_ix.find(p.mid);
      if (im < 0) deck_fail(e.line, "MAT1 " + std::to_string(p.mid) + " is not defined");
      const MatRec& mt = mats[static_cast<size_t>(im)];
      if (p.kind == 0) {
        const int isop = p.isop == 1 ? 1 : 0;
        if (isop_seen >= 0 && isop_seen != isop)
          deck_fail(e.line, "mixed PSOLID ISOP (EAS and SRI) in one deck is not supported");
        isop_seen = isop;
      }
      d.elem_ids.push_back(e.id);
      d.etype.push_back(e.type);
      for (int k = 0; k < e.nn; ++k) d.conn.push_back(node_index(e.nodes[k], e.line));
      d.eptr.push_back(static_cast<int>(d.conn.size()));
      d.E.push_back(mt.E);
      d.nu.push_back(mt.nu);
      d.rho.push_back(mt.rho);
      d.thick.push_back(p.kind == 1 ? p.t : 0.0);
    }
  d.solid_form = isop_seen == 1 ? 1 : 0;
  IdIndex elem_ix;  // element IDs are looked up by PLOAD4 only
  bool any_pload = false;
  for (const auto& ck : chunks) any_pload = any_pload || !ck.pload.empty();
  if (any_pload) elem_ix.build(d.elem_ids.size(), [&](size_t i) { return d.elem_ids[i]; });
  // RBE2 then RBE3 in file order as MPC equations (expand_rbe2 / expand_rbe3,
  // src/elements/rbe_constraints.cpp:22-181) and the DOFs they activate
  // (mark_rigid_element_dofs, src/solver/analysis_support.cpp:70-95)
  d.rigid_active.assign(static_cast<size_t>(nn), 0);
  {
    auto add = [&](std::initializer_list<std::tuple<int, int, double>> terms) {
      for (const auto& [g, c, v] : terms) {
        d.rigid_node.push_back(g);
        d.rigid_dof.push_back(c);
        d.rigid_coeff.push_back(v);
      }
      d.rigid_ptr.push_back(static_cast<int>(d.rigid_node.size()));
    };
    auto pos = [&](int i) { return &d.xyz[3 * static_cast<size_t>(i)]; };
    for (int kind : {2, 3})
      for (const auto& ck : chunks)
        for (const auto& r : ck.rbes) {
          if (r.kind != kind) continue;
          const int n0 = node_index(r.node, r.line);
          const double* p0 = pos(n0);
          if (kind == 2) {
            d.rigid_active[static_cast<size_t>(n0)] |= r.comp;
            for (int gid : r.gm) {
              const int m = node_index(gid, r.line);
              d.rigid_active[static_cast<size_t>(m)] |= r.comp;
              const double* pm = pos(m);
              const double rx = pm[0] - p0[0], ry = pm[1] - p0[1], rz = pm[2] - p0[2];
              if (r.comp & 1) add({{m, 1, 1.0}, {n0, 1, -1.0}, {n0, 5, -rz}, {n0, 6, ry}});
              if (r.comp & 2) add({{m, 2, 1.0}, {n0, 2, -1.0}, {n0, 4, rz}, {n0, 6, -rx}});
              if (r.comp & 4) add({{m, 3, 1.0}, {n0, 3, -1.0}, {n0, 4, -ry}, {n0, 5, rx}});
              for (int c = 4; c <= 6; ++c)
                if (r.comp & (1 << (c - 1))) add({{m, c, 1.0}, {n0, c, -1.0}});
            }
            continue;
          }
          d.rigid_active[static_cast<size_t>(n0)] |= r.comp;
          double wt = 0.0, wr = 0.0;
          for (const auto& g : r.groups) {
            for (int gid : g.nodes) d.rigid_active[static_cast<size_t>(node_index(gid, r.line))] |= g.comp;
            const double w = g.w * static_cast<double>(g.nodes.size());
            wt += w;
            if (g.comp & 0x38) wr += w;
          }
          if (wt < 1e-30) continue;
          for (int c = 1; c <= 6; ++c) {
            if (!(r.comp & (1 << (c - 1)))) continue;
            if (c >= 4 && wr < 1e-30) continue;
            const size_t t0 = d.rigid_node.size();
            auto term = [&](int g, int cc, double v) {
              d.rigid_node.push_back(g);
              d.rigid_dof.push_back(cc);
              d.rigid_coeff.push_back(v);
            };
            term(n0, c, c <= 3 ? wt : wr);
            for (const auto& g : r.groups) {
              if (c >= 4 && !(g.comp & (1 << (c - 1)))) continue;
              const bool rot = (g.comp & 0x38) != 0;
              for (int gid : g.nodes) {
                const int j = node_index(gid, r.line);
                term(j, c, -g.w);
                if (c >= 4 || !rot) continue;
                const double* pj = pos(j);
                const double rx = pj[0] - p0[0], ry = pj[1] - p0[1], rz = pj[2] - p0[2];
                // lever terms of (omega_j x r_j)[c - 1]
                if (c == 1) {
                  if (g.comp & 16) term(j, 5, -g.w * rz);
                  if (g.comp & 32) term(j, 6, g.w * ry);
                } else if (c == 2) {
                  if (g.comp & 32) term(j, 6, -g.w * rx);
                  if (g.comp & 8) term(j, 4, g.w * rz);
                } else {
                  if (g.comp & 8) term(j, 4, -g.w * ry);
                  if (g.comp & 16) term(j, 5, g.w * rx);
                }
              }
            }
            if (d.rigid_node.size() > t0) d.rigid_ptr.push_back(static_cast<int>(d.rigid_node.size()));
          }
        }
  }
  // model.mpcs: MPC cards in file order, each MPCADD appending relabelled
  // copies of the equations listed so far (process_mpcadd, 2070-2095)
  std::vector<const MpcRec*> mpcs;
  std::vector<int> mpc_sid;
  for (const auto& ck : chunks)
    for (const auto& r : ck.mpcs) {
      if (!r.add) {
        mpcs.push_back(&r);
        mpc_sid.push_back(r.sid);
        continue;
      }
      for (int src : r.srcs) {
        if (src == r.sid) continue;
        const size_t n0 = mpcs.size();
        for (size_t k = 0; k < n0; ++k)
          if (mpc_sid[k] == src) {
            mpcs.push_back(mpcs[k]);
            mpc_sid.push_back(r.sid);
          }
      }
    }
  // subcases: SPC set, FORCE/MOMENT, PLOAD4, GRAV of the LOAD set, MPC set
  for (auto& s : d.subs) {
    if (s.mpc != 0)
      for (size_t k = 0; k < mpcs.size(); ++k) {
        if (mpc_sid[k] != s.mpc) continue;
        for (const auto& t : mpcs[k]->terms) {
          s.mpc_node.push_back(node_index(t.node, mpcs[k]->line));
          s.mpc_dof.push_back(t.dof);
          s.mpc_coeff.push_back(t.coeff);
        }
        s.mpc_ptr.push_back(static_cast<int>(s.mpc_node.size()));
      }
    std::vector<int> mask(static_cast<size_t>(nn), 0);
    s.nodal.assign(6 * static_cast<size_t>(nn), 0.0);
    for (const auto& ck : chunks) {
      for (const auto& r : ck.spc)
        if (r.sid == s.spc) mask[static_cast<size_t>(node_index(r.node, 0))] |= r.mask;
      for (const auto& r : ck.nodal)
        if (r.sid == s.load) {
          const size_t i = static_cast<size_t>(node_index(r.node, 0));
          double v[3] = {r.v[0], r.v[1], r.v[2]};
          const auto fit = r.cid != 0 ? frames.find(r.cid) : frames.end();
          if (fit != frames.end()) {  // a CID without a frame leaves the vector as given
            double R[3][3];
            rotation(fit->second, &d.xyz[3 * i], R);
            for (int a = 0; a < 3; ++a) v[a] = R[a][0] * r.v[0] + R[a][1] * r.v[1] + R[a][2] * r.v[2];
          }
          for (int k = 0; k < 3; ++k) s.nodal[6 * i + r.off + k] += v[k];
        }
      for (const auto& r : ck.grav)
        if (r.sid == s.load)
          for (int k = 0; k < 3; ++k) s.grav[k] += r.a[k];
      for (const auto& r : ck.pload)
        if (r.sid == s.load) {
          const int ie = elem_ix.find(r.eid);
          if (ie < 0)
            deck_fail(r.line, "PLOAD4 element " + std::to_string(r.eid) + " is not defined");
          s.pl_elem.push_back(ie);
          s.pl_g1.push_back(r.g1 < 0 ? -1 : node_index(r.g1, r.line));
          s.pl_g34.push_back(r.g34 < 0 ? -1 : node_index(r.g34, r.line));
          for (int k = 0; k < 4; ++k) s.pl_p.push_back(r.p[k]);
        }
    }
    for (int i = 0; i < nn; ++i)
      if (mask[static_cast<size_t>(i)]) {
        s.spc_nodes.push_back(i);
        s.spc_masks.push_back(mask[static_cast<size_t>(i)]);
      }
  }
  d.unsupported.assign(unsupported.begin(), unsupported.end());
}

}  // namespace
}  // namespace vbh

struct vbh_deck : vbh::Deck {};

extern "C" {

int vbh_deck_parse(const char* text, long long len, vbh_deck** out) {
  return vbh::guarded([&] {
    if (!text || len < 0 || !out) vbh::throw_error(VBH_ERR_ARG, "vbh_deck_parse: bad argument");
    auto d = std::make_unique<vbh_deck>();
    try {
      vbh::parse_text(std::string_view(text, static_cast<size_t>(len)), *d);
    } catch (const vbh::DeckError& e) {
      vbh::throw_error(VBH_ERR_SOLVER, e.msg);
    }
    *out = d.release();
  });
}

int vbh_deck_read(const char* path, vbh_deck** out) {
  return vbh::guarded([&] {
    if (!path || !out) vbh::throw_error(VBH_ERR_ARG, "vbh_deck_read: bad argument");
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) vbh::throw_error(VBH_ERR_SOLVER, std::string("Cannot open BDF file: ") + path);
    const std::streamoff size = f.tellg();
    std::string text(static_cast<size_t>(size > 0 ? size : 0), '\0');
    f.seekg(0);
    if (size > 0 && !f.read(text.data(), size))
      vbh::throw_error(VBH_ERR_SOLVER, std::string("Cannot read BDF file: ") + path);
    auto d = std::make_unique<vbh_deck>();
    try {
      vbh::parse_text(text, *d);
    } catch (const vbh::DeckError& e) {
      vbh::throw_error(VBH_ERR_SOLVER, e.msg);
    }
    *out = d.release();
  });
}

void vbh_deck_free(vbh_deck* d) { delete d; }

int vbh_deck_info(const vbh_deck* d, vbh_deck_sizes* s) {
  if (!d || !s) {
    vbh::set_error("vbh_deck_info: null argument");
    return VBH_ERR_ARG;
  }
  s->nnodes = static_cast<int>(d->node_ids.size());
  s->nelem = static_cast<int>(d->elem_ids.size());
  s->nconn = static_cast<int>(d->conn.size());
  s->nsubcase = static_cast<int>(d->subs.size());
  s->solid_form = d->solid_form;
  s->shell_form = d->shell_form;
  s->nunsupported = static_cast<int>(d->unsupported.size());
  s->nrigid_mpc = static_cast<int>(d->rigid_ptr.size()) - 1;
  s->nrigid_terms = static_cast<int>(d->rigid_node.size());
  return VBH_OK;
}

int vbh_deck_mesh(const vbh_deck* d, int* node_ids, double* xyz, int* elem_ids, int* etype,
                  int* eptr, int* conn, double* E, double* nu, double* rho, double* thickness) {
  if (!d) {
    vbh::set_error("vbh_deck_mesh: null deck");
    return VBH_ERR_ARG;
  }
  auto cp = [](auto* dst, const auto& v) {
    if (dst) std::copy(v.begin(), v.end(), dst);
  };
  cp(node_ids, d->node_ids);
  cp(xyz, d->xyz);
  cp(elem_ids, d->elem_ids);
  cp(etype, d->etype);
  cp(eptr, d->eptr);
  cp(conn, d->conn);
  cp(E, d->E);
  cp(nu, d->nu);
  cp(rho, d->rho);
  cp(thickness, d->thick);
  return VBH_OK;
}

int vbh_deck_subcase(const vbh_deck* d, int k, vbh_deck_subcase_info* s) {
  if (!d || !s || k < 0 || k >= static_cast<int>(d->subs.size())) {
    vbh::set_error("vbh_deck_subcase: bad argument");
    return VBH_ERR_ARG;
  }
  const auto& u = d->subs[static_cast<size_t>(k)];
  s->id = u.id;
  s->load_set = u.load;
  s->spc_set = u.spc;
  s->nspc = static_cast<int>(u.spc_nodes.size());
  s->npload4 = static_cast<int>(u.pl_elem.size());
  s->disp_print = u.disp;
  s->stress_print = u.stress;
  s->mpc_set = u.mpc;
  s->nmpc = static_cast<int>(u.mpc_ptr.size()) - 1;
  s->nmpc_terms = static_cast<int>(u.mpc_node.size());
  for (int i = 0; i < 3; ++i) s->grav[i] = u.grav[i];
  return VBH_OK;
}

int vbh_deck_subcase_data(const vbh_deck* d, int k, int* spc_nodes, int* spc_masks,
                          double* nodal, int* pl_elem, int* pl_g1, int* pl_g34, double* pl_p) {
  if (!d || k < 0 || k >= static_cast<int>(d->subs.size())) {
    vbh::set_error("vbh_deck_subcase_data: bad argument");
    return VBH_ERR_ARG;
  }
  const auto& u = d->subs[static_cast<size_t>(k)];
  auto cp = [](auto* dst, const auto& v) {
    if (dst) std::copy(v.begin(), v.end(), dst);
  };
  cp(spc_nodes, u.spc_nodes);
  cp(spc_masks, u.spc_masks);
  cp(nodal, u.nodal);
  cp(pl_elem, u.pl_elem);
  cp(pl_g1, u.pl_g1);
  cp(pl_g34, u.pl_g34);
  cp(pl_p, u.pl_p);
  return VBH_OK;
}

int vbh_deck_subcase_mpc(const vbh_deck* d, int k, int* mpc_ptr, int* term_node, int* term_dof,
                         double* term_coeff) {
  if (!d || k < 0 || k >= static_cast<int>(d->subs.size())) {
    vbh::set_error("vbh_deck_subcase_mpc: bad argument");
    return VBH_ERR_ARG;
  }
  const auto& u = d->subs[static_cast<size_t>(k)];
  auto cp = [](auto* dst, const auto& v) {
    if (dst) std::copy(v.begin(), v.end(), dst);
  };
  cp(mpc_ptr, u.mpc_ptr);
  cp(term_node, u.mpc_node);
  cp(term_dof, u.mpc_dof);
  cp(term_coeff, u.mpc_coeff);
  return VBH_OK;
}

int vbh_deck_rigid(const vbh_deck* d, int* active, int* mpc_ptr, int* term_node, int* term_dof,
                   double* term_coeff) {
  if (!d) {
    vbh::set_error("vbh_deck_rigid: null deck");
    return VBH_ERR_ARG;
  }
  auto cp = [](auto* dst, const auto& v) {
    if (dst) std::copy(v.begin(), v.end(), dst);
  };
  cp(active, d->rigid_active);
  cp(mpc_ptr, d->rigid_ptr);
  cp(term_node, d->rigid_node);
  cp(term_dof, d->rigid_dof);
  cp(term_coeff, d->rigid_coeff);
  return VBH_OK;
}

int vbh_deck_frames(const vbh_deck* d, int* cd, double* rot) {
  if (!d) {
    vbh::set_error("vbh_deck_frames: null deck");
    return VBH_ERR_ARG;
  }
  if (cd) std::copy(d->cd.begin(), d->cd.end(), cd);
  if (rot) std::copy(d->cd_T.begin(), d->cd_T.end(), rot);
  return VBH_OK;
}

const char* vbh_deck_unsupported(const vbh_deck* d, int i) {
  if (!d || i < 0 || i >= static_cast<int>(d->unsupported.size())) return nullptr;
  return d->unsupported[static_cast<size_t>(i)].c_str();
}

}  // extern "C"
