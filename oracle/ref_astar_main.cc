// Checker driver for oracle/_ref (test infrastructure, never shipped): builds
// the reference's own phase-3 planar corpus (src/benchmark/planar_corpus.cc,
// BuildPhase3CandidateCorpus) from the KiCad fixture text given on the command
// line, routes every case with the reference's CPU A* (RouteWithCpuAStar,
// src/routing/cpu_astar.cc) and prints one JSON object per case: name, status,
// cost, the search telemetry and the lattice path.  Compiled together with the
// reference's sources by oracle/Makefile (target `ref`); the output pins the
// restated A* of oracle/v1.c case by case.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <variant>

#include "apgar/benchmark/planar_corpus.h"
#include "apgar/routing/cpu_astar.h"

namespace {

std::string Escape(const std::string& s) {
  std::string out;
  for (char c : s) {
    if (c == '"' || c == '\\') out.push_back('\\');
    if (static_cast<unsigned char>(c) < 0x20) continue;
    out.push_back(c);
  }
  return out;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <kicad fixture>\n", argv[0]);
    return 2;
  }
  std::ifstream in(argv[1], std::ios::binary);
  if (!in) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  std::stringstream text;
  text << in.rdbuf();
  const std::string fixture = text.str();
  auto corpus = apgar::benchmark::BuildPhase3CandidateCorpus(fixture);
  if (std::holds_alternative<std::string>(corpus)) {
    std::fprintf(stderr, "corpus: %s\n", std::get<std::string>(corpus).c_str());
    return 1;
  }
  for (const auto& c : std::get<std::vector<apgar::benchmark::PlanarCorpusCase>>(corpus)) {
    const auto result = apgar::routing::RouteWithCpuAStar(c.board, c.compiled_board, c.request);
    std::ostringstream o;
    o << "{\"name\": \"" << Escape(c.name) << "\"";
    const apgar::routing::CpuRouteTelemetry* t = nullptr;
    if (std::holds_alternative<apgar::routing::CpuRoute>(result)) {
      const auto& r = std::get<apgar::routing::CpuRoute>(result);
      t = &r.telemetry;
      o << ", \"status\": \"routed\", \"cost\": " << r.total_cost << ", \"path\": [";
      for (size_t i = 0; i < r.lattice_path.size(); ++i)
        o << (i ? ", " : "") << "[" << r.lattice_path[i].x << ", " << r.lattice_path[i].y << "]";
      o << "]";
    } else {
      const auto& f = std::get<apgar::routing::RouteFailure>(result);
      o << ", \"status\": \"failed\", \"failure_code\": " << static_cast<int>(f.code) << ", \"detail\": \""
        << Escape(f.detail) << "\"";
      if (f.telemetry) t = &*f.telemetry;
    }
    if (t)
      o << ", \"work_units\": " << t->work_units << ", \"queue_pops\": " << t->queue_pops
        << ", \"expanded_states\": " << t->expanded_states << ", \"attempted_relaxations\": "
        << t->attempted_relaxations << ", \"accepted_relaxations\": " << t->accepted_relaxations;
    o << "}";
    std::cout << o.str() << "\n";
  }
  return 0;
}
