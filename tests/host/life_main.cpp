// Drives the generated host class of Life the way the reference's examples/Life/main.cpp
// drives its generated class: construct, init, set the cells, proceed, read back.
// Usage: life_main <machine.so> <initial cells file> <generations> <final cells file>
// The cells files hold om_memory_size() int32 values, x-fastest.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "Life.hpp"

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s machine.so init.bin generations out.bin\n", argv[0]);
    return 2;
  }
  try {
    Life life(argv[1]);
    life.init();
    std::vector<int> cells(life.om_memory_size());
    FILE* f = std::fopen(argv[2], "rb");
    if (!f || std::fread(cells.data(), sizeof(int), cells.size(), f) != cells.size()) {
      std::fprintf(stderr, "cannot read %s\n", argv[2]);
      return 1;
    }
    std::fclose(f);
    // live cells through the elemental accessor, as examples/Life/main.cpp sets them
    const int W = life.om_size_0(), H = life.om_size_1();
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x)
        if (cells[x + W * y]) life.cell(x, y) = 1;
    const int gens = std::atoi(argv[3]);
    for (int g = 0; g < gens; ++g) life.proceed();
    const std::vector<int> out = life.cell();   // brought back from the device lazily
    f = std::fopen(argv[4], "wb");
    if (!f || std::fwrite(out.data(), sizeof(int), out.size(), f) != out.size()) return 1;
    std::fclose(f);
    std::printf("%d %d\n", life.population(), life.generation());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
