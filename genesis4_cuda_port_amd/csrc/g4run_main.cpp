// g4run - runs a GENESIS 1.3 input deck on one GPU through g4h_deck_run.
//
//   g4run deck.in [-o out_dir] [-d device]
//
// Records are written as .npy arrays under out_dir (default: the deck's
// name with ".out"), named like the reference's output groups.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/g4hip.h"

int main(int argc, char** argv) {
  std::string deck, out;
  int device = 0;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "-o") && i + 1 < argc) {
      out = argv[++i];
    } else if (!std::strcmp(argv[i], "-d") && i + 1 < argc) {
      device = std::atoi(argv[++i]);
    } else if (argv[i][0] == '-' || !deck.empty()) {
      std::fprintf(stderr, "usage: %s deck.in [-o out_dir] [-d device]\n", argv[0]);
      return 2;
    } else {
      deck = argv[i];
    }
  }
  if (deck.empty()) {
    std::fprintf(stderr, "usage: %s deck.in [-o out_dir] [-d device]\n", argv[0]);
    return 2;
  }
  if (out.empty()) {
    out = deck;
    const size_t dot = out.find_last_of('.');
    if (dot != std::string::npos && out.find('/', dot) == std::string::npos) out.resize(dot);
    out += ".out";
  }
  int nrec = 0;
  const g4h_status st = g4h_deck_run(deck.c_str(), out.c_str(), device, &nrec);
  if (st != G4H_OK) {
    std::fprintf(stderr, "g4run: %s\n", g4h_last_error());
    return 1;
  }
  std::printf("g4run: %d records written to %s\n", nrec, out.c_str());
  return 0;
}
