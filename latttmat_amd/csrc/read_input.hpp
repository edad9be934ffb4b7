// read_input.hpp -- the tmat input-file parser shared by the tmat and
// dumpconfig programs (TMat/read_input.cpp:14-92, used by TMat/main.cpp and
// DumpConfig/main.cpp alike).
#pragma once

#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>

namespace tmcli {

struct ReadInput {
  char model = 0;
  int dim = 0, axialdir = 0, neigs = 0;
  int savevec[2] = {0, 0};
  std::string outprefix;
  int size = 0, size1 = 0, size2 = 0;
  double betaJ = 0, kappa = 0, kappa2 = 0, h = 0;

  // TMat/read_input.cpp:14-92: every value is read after the next '=';
  // key names are not checked.
  int read(const char *inputf)
  {
    std::ifstream in(inputf);
    if (!in) {
      std::cout << "Can't open " << inputf << " for input." << std::endl;
      return 2;
    }
    std::cout << "Reading input from file " << inputf << std::endl;
    auto next = [&]() {
      in.ignore(std::numeric_limits<std::streamsize>::max(), '=');
    };
    next(); in >> model;
    next(); in >> dim;
    next(); in >> axialdir;
    next(); in >> neigs;
    next(); in >> savevec[0] >> savevec[1];
    next(); in >> outprefix;
    next();
    if (dim == 2) {
      in >> size;
      size1 = size;
    } else if (dim == 3) {
      in >> size1 >> size2;
      size = size1 << 16 | size2;
    } else {
      std::cout << "Invalid dim " << dim << std::endl;
      return 4;
    }
    next(); in >> betaJ;
    next();
    std::string kline;
    std::getline(in, kline);
    std::istringstream ks(kline);
    ks >> kappa;
    if (!(ks >> kappa2)) kappa2 = 0.0;
    next(); in >> h;
    int error = 0;
    if (in.fail()) {
      std::cout << "Error reading input file " << inputf << std::endl;
      error = 3;
    }
    std::cout << "     MODEL : " << dim << "D_" << model << axialdir << "\n";
    std::cout << "Save & flg : " << outprefix << " " << savevec[0] << " " << savevec[1] << "\n";
    if (dim == 2) std::cout << "      size : " << size << "\n";
    else std::cout << "      size : " << size1 << " " << size2 << "\n";
    std::cout << "     betaJ : " << betaJ << "\n";
    std::cout << "     kappa : " << kappa << std::endl;
    if (model == 'B') std::cout << "    kappa2 : " << kappa2 << std::endl;
    std::cout << "         h : " << h << std::endl;
    if (model != 'A' && model != 'B' && model != 'D') {
      std::cout << "Invalid model " << model << std::endl;
      error = 4;
    } else if (size < 4) {
      std::cout << "Invalid size (<4) " << size << std::endl;
      error = 4;
    }
    return error;
  }
};

}  // namespace tmcli
