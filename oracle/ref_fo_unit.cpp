// Test infrastructure only.  Compiles the reference's freezeout helpers
// (freezeout/FreezeOut.cpp, which itself includes the Cornelius 1.3 source, and
// freezeout/Memory.cpp) as one translation unit, the same way the reference's
// driver pulls them in textually (rhic/src/HydroPlugin.cpp:21-22).  The files
// are read from the reference tree by oracle/ref.mk; nothing is copied.
#include <cstdlib>
#include <cstdio>
#include <vector>
#include <fstream>
#include "FreezeOut.cpp"
#include "Memory.cpp"
