// placeholder driver (replaced below)
int main() { return 0; }
