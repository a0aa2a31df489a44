/* config.h — stands in for the header that the reference's configure script generates and
   its test programs include after meep.hpp (tests/known_results.cpp:25).  Those programs read
   nothing from it when built against this library, so it defines nothing. */
