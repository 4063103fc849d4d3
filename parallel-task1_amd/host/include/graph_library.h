// include/graph_library.h -- the umbrella header the reference's drivers
// include by this path (sample/sssp.cpp:1, sample/solve.cpp:1). Compiled with
// -I <this directory's parent> it resolves to the MI355X library's host side:
// the standard headers and `using namespace std` the reference's umbrella
// supplies to its drivers (through include/graph_data_structures/graph.h and
// algorithm/sssp/sssp.h:19), then the interface mirror.
#pragma once

#include <omp.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "../graph_library.h"

using namespace std;
