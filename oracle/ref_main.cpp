// main() for the reference C model (R/C_Model_ForceEval/MD_Evaluation_Model.cpp), which
// has no entry point of its own.  The model reads its input from the relative path
// "C:/Users/Ethan/.../MatlabScripts/ar_gas.pdb" (:30-32); the test runs this binary in
// a scratch directory holding a generated PDB at that path.
int MD_Evaluation_Model();
int main() { return MD_Evaluation_Model() == 1 ? 0 : 1; }
