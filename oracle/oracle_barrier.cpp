// ORACLE — test infrastructure only (see oracle_interval.hpp).
