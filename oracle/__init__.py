"""CPU oracle of the diffuse path (test infrastructure only; see oracle.cpp)."""
