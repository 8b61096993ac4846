"""CPU checker (test infrastructure only): FP64 restatement of the reference's
acoustic solver and gradient, plus the build recipe of the reference's own C++
core in oracle/_ref.  Never imported by the product package."""
