"""The standard gate library the parser preloads (the role of qelib1.inc, qfasm_lexer.py:56-58).

Gates with a native class (parser.NATIVE) only need their signature here. Every other library gate
is expanded through the definition below; these are the OpenQASM 2.0 standard-library
decompositions (qelib1.inc of the OpenQASM 2.0 specification, as shipped with pyquafu). They must be
kept exactly: the decomposition fixes the global phase of the simulated state (u3 reaches the engine
as rz, ry, rz, not as the U3 matrix).
"""

# (name, number of parameters, number of qubits) of the natively mapped gates
NATIVE_SIGNATURES = [
    ("x", 0, 1), ("y", 0, 1), ("z", 0, 1), ("h", 0, 1), ("s", 0, 1), ("sdg", 0, 1), ("t", 0, 1), ("tdg", 0, 1),
    ("id", 0, 1), ("sx", 0, 1), ("sy", 0, 1), ("w", 0, 1), ("sw", 0, 1), ("rx", 1, 1), ("ry", 1, 1), ("rz", 1, 1),
    ("p", 1, 1), ("delay", 1, 1), ("cx", 0, 2), ("cnot", 0, 2), ("cy", 0, 2), ("cz", 0, 2), ("cs", 0, 2),
    ("ct", 0, 2), ("swap", 0, 2), ("cp", 1, 2), ("rxx", 1, 2), ("ryy", 1, 2), ("rzz", 1, 2), ("ccx", 0, 3),
    ("toffoli", 0, 3), ("cswap", 0, 3), ("fredkin", 0, 3),
]

QELIB1 = """
gate u3(theta, phi, lambda) q { U(theta, phi, lambda) q; }
gate u2(phi, lambda) q { U(pi/2, phi, lambda) q; }
gate u1(lambda) q { U(0, 0, lambda) q; }
gate u0(gamma) q { U(0, 0, 0) q; }
gate u(theta, phi, lambda) q { U(theta, phi, lambda) q; }
gate sxdg a { s a; h a; s a; }
gate ch a, b { h b; sdg b; cx a, b; h b; t b; cx a, b; t b; h b; s b; x b; s a; }
gate crx(lambda) a, b { u1(pi/2) b; cx a, b; u3(-lambda/2, 0, 0) b; cx a, b; u3(lambda/2, -pi/2, 0) b; }
gate cry(lambda) a, b { ry(lambda/2) b; cx a, b; ry(-lambda/2) b; cx a, b; }
gate crz(lambda) a, b { rz(lambda/2) b; cx a, b; rz(-lambda/2) b; cx a, b; }
gate cu1(lambda) a, b { u1(lambda/2) a; cx a, b; u1(-lambda/2) b; cx a, b; u1(lambda/2) b; }
gate cu3(theta, phi, lambda) c, t {
  u1((lambda + phi)/2) c; u1((lambda - phi)/2) t; cx c, t;
  u3(-theta/2, 0, -(phi + lambda)/2) t; cx c, t; u3(theta/2, phi, 0) t;
}
gate csx a, b { h b; cu1(pi/2) a, b; h b; }
gate cu(theta, phi, lambda, gamma) c, t {
  p(gamma) c; p((lambda + phi)/2) c; p((lambda - phi)/2) t; cx c, t;
  u(-theta/2, 0, -(phi + lambda)/2) t; cx c, t; u(theta/2, phi, 0) t;
}
gate rccx a, b, c {
  u2(0, pi) c; u1(pi/4) c; cx b, c; u1(-pi/4) c; cx a, c; u1(pi/4) c; cx b, c; u1(-pi/4) c; u2(0, pi) c;
}
gate rc3x a, b, c, d {
  u2(0, pi) d; u1(pi/4) d; cx c, d; u1(-pi/4) d; u2(0, pi) d; cx a, d; u1(pi/4) d; cx b, d;
  u1(-pi/4) d; cx a, d; u1(pi/4) d; cx b, d; u1(-pi/4) d; u2(0, pi) d; u1(pi/4) d; cx c, d;
  u1(-pi/4) d; u2(0, pi) d;
}
gate c3x a, b, c, d {
  h d; p(pi/8) a; p(pi/8) b; p(pi/8) c; p(pi/8) d;
  cx a, b; p(-pi/8) b; cx a, b; cx b, c; p(-pi/8) c; cx a, c; p(pi/8) c; cx b, c; p(-pi/8) c;
  cx a, c; cx c, d; p(-pi/8) d; cx b, d; p(pi/8) d; cx c, d; p(-pi/8) d; cx a, d; p(pi/8) d;
  cx c, d; p(-pi/8) d; cx b, d; p(pi/8) d; cx c, d; p(-pi/8) d; cx a, d; h d;
}
gate c3sqrtx a, b, c, d {
  h d; cu1(pi/8) a, d; h d; cx a, b; h d; cu1(-pi/8) b, d; h d; cx a, b; h d; cu1(pi/8) b, d; h d;
  cx b, c; h d; cu1(-pi/8) c, d; h d; cx a, c; h d; cu1(pi/8) c, d; h d; cx b, c;
  h d; cu1(-pi/8) c, d; h d; cx a, c; h d; cu1(pi/8) c, d; h d;
}
gate c4x a, b, c, d, e {
  h e; cu1(pi/2) d, e; h e; c3x a, b, c, d; h e; cu1(-pi/2) d, e; h e; c3x a, b, c, d; c3sqrtx a, b, c, e;
}
"""
