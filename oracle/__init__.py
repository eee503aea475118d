"""CPU oracle for the acceleration-structure path. TEST INFRASTRUCTURE ONLY (see oracle.cpp)."""
